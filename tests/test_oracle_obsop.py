"""Pins of the oracle's observation operator (gen_obs_struct,
letkf_obs.F90:16-187) by what the mathematics fixes:

* bilinear interpolation in lat/lon and linear interpolation in ln p are exact
  for a field linear in lat, lon and ln p, so H of such a field is known in
  closed form -- including the closing cells (grid closed by 90 and 360,
  pro_info.F90:77-78) and the one-sided cells below the first grid latitude
  and west of the first longitude (weight 0 on the halo point, find_obs_grid
  letkf_obs.F90:234-246);
* obs above the top / below the bottom level take the top / bottom level
  value (obs_make.F90:299-313);
* centring: sum_e Yb = 0 and dep + mean_e H(x_e) = yo;
* the level search continues from the previous level of the same site
  (obs_make.F90:251-258): for levels out of order, the bracket of the earlier
  level is kept, checked against a hand-computed chord on a 4-level column.
"""
import numpy as np
import pytest

from dida_letkf import synth


def _linear_case(seed=5, k=6, nz=5):
    c = synth.obs_operator_case("c2", seed=seed, k=k, nlat=12, nlon=24, nz=nz, n_obs=300,
                                n_stations=6, n_ps=20)
    inst = c.inst
    cfg = inst.cfg
    h = cfg.halo
    rng = np.random.default_rng(seed)
    P = 9.6e4
    ce = rng.uniform(-500.0, 500.0, k)           # member ps offsets, uniform in space
    ps_e = P + ce
    hyai, hybi = np.asarray(cfg.hyai), np.asarray(cfg.hybi)
    lev = hyai[None, :] * (cfg.p0 - cfg.ptop) + hybi[None, :] * (ps_e[:, None] - cfg.ptop) + cfg.ptop
    ph = 0.5 * (lev[:, :-1] + lev[:, 1:])        # [k, nz]
    beta, gamma, delta = 0.37, -0.011, 2.5
    alpha = {v: 10.0 * (v + 1) for v in range(cfg.n_var)}
    eps = rng.standard_normal(k)
    # coordinates of every array row/column, halo included: closing entries 90 / 360
    lat_ext = np.concatenate([[-95.0], inst.grid_lat.astype(np.float64), [90.0]])
    lon_ext = np.concatenate([[-5.0], inst.grid_lon.astype(np.float64), [360.0]])
    LAT = lat_ext[None, :]
    LON = lon_ext[:, None]
    vs = cfg.var_start() - 1
    inst.axb[...] = 0.0
    inst.xb[...] = 0.0
    inst.axb[:, :, 0] = P
    inst.xb[:, :, 0, :] = ce[None, None, :]
    for v in range(cfg.n_var):
        if cfg.var_lev()[v] == 1:
            continue
        for L in range(cfg.nz):
            inst.axb[:, :, vs[v] + L] = alpha[v] + beta * LAT + gamma * LON
            inst.xb[:, :, vs[v] + L, :] = delta * np.log(ph[:, L])[None, None, :] + eps[None, None, :]
    return c, dict(ps_e=ps_e, ph=ph, beta=beta, gamma=gamma, delta=delta, alpha=alpha, eps=eps,
                   g0=float(inst.grid_lat[0]), l0=float(inst.grid_lon[0]))


def _expected_H(c, par):
    inst = c.inst
    cfg = inst.cfg
    k = cfg.k
    rows = int(c.site_row_ptr[-1])
    H = np.zeros((rows, k))
    for s in range(c.site_lat.size):
        lat = max(c.site_lat[s], par["g0"])
        lon = max(c.site_lon[s], par["l0"])
        for r in range(c.site_row_ptr[s], c.site_row_ptr[s + 1]):
            v = c.row_var[r]
            if cfg.var_lev()[v] == 1:
                H[r] = par["ps_e"]
                continue
            p = c.row_ph[r]
            ph = par["ph"]
            lp = np.where(p <= ph[:, 0], np.log(ph[:, 0]),
                          np.where(p > ph[:, -1], np.log(ph[:, -1]), np.log(p)))
            H[r] = par["alpha"][v] + par["beta"] * lat + par["gamma"] * lon + par["delta"] * lp + par["eps"]
    return H


def test_linear_field_is_reproduced(orc):
    c, par = _linear_case()
    yb, dep = orc.obs_operator(c)
    H = _expected_H(c, par)
    mean = H.mean(axis=1)
    scale = np.abs(H).max()
    assert np.abs(yb - (H - mean[:, None])).max() < 1e-12 * scale
    assert np.abs(dep - (c.yo - mean)).max() < 1e-12 * scale
    # the closing cells and one-sided cells were exercised
    assert (c.site_lat > 89.5).any() and (c.site_lon > 359.5).any()
    assert (c.site_lat < -89.5).any() and (c.site_lon < 0.5).any()


def test_out_of_range_levels_clamp(orc):
    c, par = _linear_case(seed=8)
    ph = par["ph"]
    # obs above the top level of every member and below the bottom one
    c.row_ph[0] = 0.5 * ph[:, 0].min()
    c.row_ph[1] = 1.5 * ph[:, -1].max()
    c.row_var[0] = c.row_var[1] = 1
    yb, dep = orc.obs_operator(c)
    H = _expected_H(c, par)
    for r in (0, 1):
        mean = H[r].mean()
        assert np.abs(yb[r] - (H[r] - mean)).max() < 1e-12 * np.abs(H[r]).max()


def test_centring_invariants(orc):
    c = synth.obs_operator_case("c2", seed=11, k=20)
    yb, dep = orc.obs_operator(c)
    mag = np.abs(yb).max(axis=1) + 1e5 * (c.row_var == 0)
    assert np.all(np.abs(yb.sum(axis=1)) <= 1e-13 * c.inst.cfg.k * np.maximum(mag, 1.0))
    # dep = yo - mean H: H = yb + mean, so yo - dep is the mean and yb + (yo - dep) = H
    c2 = synth.obs_operator_case("c2", seed=11, k=20)
    c2.yo = c2.yo + 3.0
    _, dep2 = orc.obs_operator(c2)
    assert np.allclose(dep2 - dep, 3.0, rtol=0, atol=1e-9)


def test_level_search_continues_across_levels(orc):
    """Pure-pressure levels (hybi = 0): mid-level pressures 20, 40, 60, 80 kPa.
    A site with levels (70 kPa, 30 kPa): the first lands in bracket (3, 4);
    the search does not restart, so the second level keeps bracket (3, 4) and
    is extrapolated along that chord of the field f = (ln p)^2."""
    c = synth.obs_operator_case("c2", seed=2, k=3, nlat=8, nlon=16, nz=4, n_obs=1,
                                n_stations=0, n_ps=0)
    inst = c.inst
    cfg = inst.cfg
    ptop, p0 = 0.0, 1.0e5
    # interfaces 10, 30, 50, 70, 90 kPa -> mid levels 20, 40, 60, 80 kPa
    cfg.hyai = [0.1, 0.3, 0.5, 0.7, 0.9]
    cfg.hybi = [0.0] * 5
    cfg.p0, cfg.ptop = p0, ptop
    mids = np.array([2e4, 4e4, 6e4, 8e4])
    vs = cfg.var_start() - 1
    inst.xb[...] = 0.0
    inst.axb[...] = 0.0
    inst.axb[:, :, 0] = 1.0e5
    for L in range(4):
        inst.axb[:, :, vs[1] + L] = np.log(mids[L]) ** 2
    c.site_lat = np.array([10.0])
    c.site_lon = np.array([100.0])
    c.site_row_ptr = np.array([0, 2], np.int32)
    c.row_var = np.array([1, 1], np.int32)
    c.row_ph = np.array([7.0e4, 3.0e4])
    c.yo = np.zeros(2)
    yb, dep = orc.obs_operator(c)
    H = -dep                                     # yo = 0, yb = 0 (no spread)
    f = lambda p: np.log(p) ** 2
    a, b = np.log(6e4), np.log(8e4)

    def chord(x):
        t3 = (b - np.log(x)) / (b - a)
        return f(6e4) * t3 + f(8e4) * (1 - t3)

    assert abs(H[0] - chord(7e4)) < 1e-12 * abs(H[0])
    assert abs(H[1] - chord(3e4)) < 1e-12 * abs(H[1])
    # a restarted search would use bracket (1, 2) and give a different value
    a2, b2 = np.log(2e4), np.log(4e4)
    t3 = (b2 - np.log(3e4)) / (b2 - a2)
    assert abs(H[1] - (f(2e4) * t3 + f(4e4) * (1 - t3))) > 1e-3


def test_requires_halo_and_hybrid(orc):
    c = synth.obs_operator_case("c2", seed=4, k=4, nlat=8, nlon=16, nz=3, n_obs=10,
                                n_stations=1, n_ps=1)
    c.inst.cfg.hyai = None
    with pytest.raises(ValueError):
        orc.obs_operator(c)
