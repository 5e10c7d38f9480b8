"""GPU parity of the observation operator (dida_letkf_obs_operator_device)
against the oracle's restatement of gen_obs_struct, and an end-to-end cycle
H(x) -> analysis with every input on the device.  FP64 bar: 1e-10 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("k,nz,n_obs", [(20, 6, 3000), (64, 8, 5000), (128, 5, 2000), (3, 4, 50)])
def test_obs_operator_matches_oracle(k, nz, n_obs):
    import orc
    from dida_letkf import synth, engine
    c = synth.obs_operator_case("c2", seed=100 + k, k=k, nlat=18, nlon=36, nz=nz, n_obs=n_obs,
                                n_stations=25, n_ps=60)
    yb_o, dep_o = orc.obs_operator(c)
    yb_g, dep_g = engine.obs_operator(c)
    # Yb is a deviation from the mean of values up to ~1e5 (ps rows): scale by the values
    H_o = yb_o + (c.yo - dep_o)[:, None]
    H_g = yb_g + (c.yo - dep_g)[:, None]
    print(k, _rel(H_g, H_o), _rel(yb_g, yb_o), _rel(dep_g, dep_o))
    assert _rel(H_g, H_o) < TOL
    for rows in (c.row_var == 0, c.row_var != 0):      # ps rows and the 3-D rows separately
        assert np.abs(yb_g[rows] - yb_o[rows]).max() <= TOL * np.abs(H_o[rows]).max()
        assert np.abs(dep_g[rows] - dep_o[rows]).max() <= TOL * np.abs(H_o[rows]).max()


def test_operator_then_analysis_on_device():
    """H(x) on the device feeding the analysis equals the oracle's H feeding
    the oracle's analysis (letkf_run: gen_obs_struct then the compute loop)."""
    import torch
    import orc
    from dida_letkf import synth, engine
    c = synth.obs_operator_case("c2", seed=7, k=32, nlat=16, nlon=32, nz=5, n_obs=4000,
                                n_stations=20, n_ps=50)
    inst = c.inst
    cfg = inst.cfg
    cfg.local_dist = 900.0
    yb_o, dep_o = orc.obs_operator(c)
    inst.obs = synth.Obs(site_lat=c.site_lat, site_lon=c.site_lon, site_row_ptr=c.site_row_ptr,
                         yb=yb_o, dep=dep_o, rinv=np.full(yb_o.shape[0], 0.5),
                         vcoord=np.log(c.row_ph), row_var=c.row_var)
    ref = orc.letkf_compute(inst)
    dev = torch.device("cuda:0")
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).to(dev)
    xb, axb = t(inst.xb), t(inst.axb)
    la, lo, rp = t(c.site_lat), t(c.site_lon), t(c.site_row_ptr, np.int32)
    rv, rph, yo = t(c.row_var, np.int32), t(c.row_ph), t(c.yo)
    rinv, vco = t(inst.obs.rinv), t(np.log(c.row_ph))
    nr = int(c.site_row_ptr[-1])
    yb = torch.empty((nr, cfg.k), dtype=torch.float64, device=dev)
    dep = torch.empty(nr, dtype=torch.float64, device=dev)
    xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
    with engine.LetkfEngine(cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.obs_operator_device(xb.data_ptr(), axb.data_ptr(), c.site_lat.size, la.data_ptr(),
                                lo.data_ptr(), rp.data_ptr(), nr, rv.data_ptr(), rph.data_ptr(),
                                yo.data_ptr(), yb.data_ptr(), dep.data_ptr())
        eng.set_obs_device(c.site_lat.size, la.data_ptr(), lo.data_ptr(), rp.data_ptr(), nr,
                           yb.data_ptr(), dep.data_ptr(), rinv.data_ptr(), vco.data_ptr())
        eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr())
        torch.cuda.synchronize()
    h = cfg.halo
    g = (xa + axa[..., None]).cpu().numpy()[h:-h, h:-h]
    r = (ref["xa"] + ref["axa"][..., None])[h:-h, h:-h]
    assert _rel(g, r) < TOL


def test_operator_rejects_missing_halo():
    from dida_letkf import synth, engine
    c = synth.obs_operator_case("c2", seed=9, k=8, nlat=8, nlon=16, nz=3, n_obs=20,
                                n_stations=1, n_ps=2)
    cfg = c.inst.cfg
    cfg.halo = 0
    with engine.LetkfEngine(cfg) as eng:
        eng.set_grid(c.inst.grid_lat, c.inst.grid_lon)
        with pytest.raises(engine.DidaError):
            eng.obs_operator_device(1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1)


def test_fill_halo_matches_host_rule():
    import torch
    from dida_letkf import synth, engine
    inst = synth.generate("c1", seed=3, nlat=6, nlon=10, nz=2, n_obs=10, halo=2)
    cfg = inst.cfg
    rng = np.random.default_rng(0)
    x = rng.standard_normal(inst.xb.shape)
    ax = rng.standard_normal(inst.axb.shape)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x.copy()).to(dev)
    ad = torch.from_numpy(ax.copy()).to(dev)
    with engine.LetkfEngine(cfg) as eng:
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.fill_halo_device(xd.data_ptr(), ad.data_ptr())
        torch.cuda.synchronize()
    assert np.array_equal(xd.cpu().numpy(), synth.fill_halo(x.copy(), 2))
    assert np.array_equal(ad.cpu().numpy(), synth.fill_halo(ax.copy(), 2))


def test_obs_records_through_operator_and_analysis():
    """Obs ingestion end to end: the committed record fixtures (profiles of
    u, v, T and surface ps, read_obs_mod.F90:338-449) become rows, the
    observation operator runs on the device and the analysis follows; both
    agree with the oracle's operator and analysis on the same rows."""
    import dataclasses
    import os
    import torch
    import orc
    from dida_letkf import synth, engine, obs_records
    here = os.path.dirname(os.path.abspath(__file__))
    parts = []
    for name, var_index in (("profiles", [1, 2, 3]), ("surface", [0])):
        rec = obs_records.read_records(os.path.join(here, "golden", f"obs_records_{name}.bin"))
        parts.append(obs_records.to_rows(rec, var_index))
    # one obs set: the surface sites after the profile sites
    off = parts[0]["site_row_ptr"][-1]
    rows = {key: np.concatenate([parts[0][key], parts[1][key]]) for key in
            ("site_lat", "site_lon", "row_var", "row_ph", "yo", "rinv")}
    rows["site_row_ptr"] = np.concatenate([parts[0]["site_row_ptr"],
                                           parts[1]["site_row_ptr"][1:] + off]).astype(np.int32)
    c = synth.obs_operator_case("c2", seed=11, k=24, nlat=18, nlon=36, nz=5, n_obs=10,
                                n_stations=2, n_ps=2)
    c = dataclasses.replace(c, site_lat=rows["site_lat"], site_lon=rows["site_lon"],
                            site_row_ptr=rows["site_row_ptr"], row_var=rows["row_var"],
                            row_ph=rows["row_ph"], yo=rows["yo"])
    inst = c.inst
    cfg = inst.cfg
    cfg.local_dist = 2500.0
    yb_o, dep_o = orc.obs_operator(c)
    vco = np.log(c.row_ph)
    inst.obs = synth.Obs(site_lat=c.site_lat, site_lon=c.site_lon, site_row_ptr=c.site_row_ptr,
                         yb=yb_o, dep=dep_o, rinv=rows["rinv"], vcoord=vco, row_var=c.row_var)
    ref = orc.letkf_compute(inst, want_trans=True)
    assert ref["nobsl"].max() > cfg.k
    dev = torch.device("cuda:0")
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).to(dev)
    xb, axb = t(inst.xb), t(inst.axb)
    la, lo, rp = t(c.site_lat), t(c.site_lon), t(c.site_row_ptr, np.int32)
    rv, rph, yo = t(c.row_var, np.int32), t(c.row_ph), t(c.yo)
    rinv, vc = t(rows["rinv"]), t(vco)
    nr = int(c.site_row_ptr[-1])
    yb = torch.empty((nr, cfg.k), dtype=torch.float64, device=dev)
    dep = torch.empty(nr, dtype=torch.float64, device=dev)
    xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
    with engine.LetkfEngine(cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.obs_operator_device(xb.data_ptr(), axb.data_ptr(), c.site_lat.size, la.data_ptr(),
                                lo.data_ptr(), rp.data_ptr(), nr, rv.data_ptr(), rph.data_ptr(),
                                yo.data_ptr(), yb.data_ptr(), dep.data_ptr())
        torch.cuda.synchronize()
        H_o = yb_o + (c.yo - dep_o)[:, None]
        H_g = yb.cpu().numpy() + (c.yo - dep.cpu().numpy())[:, None]
        assert _rel(H_g, H_o) < TOL
        eng.set_obs_device(c.site_lat.size, la.data_ptr(), lo.data_ptr(), rp.data_ptr(), nr,
                           yb.data_ptr(), dep.data_ptr(), rinv.data_ptr(), vc.data_ptr())
        eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr())
        torch.cuda.synchronize()
    h = cfg.halo
    g = (xa + axa[..., None]).cpu().numpy()[h:-h, h:-h]
    r = (ref["xa"] + ref["axa"][..., None])[h:-h, h:-h]
    print("records -> operator -> analysis", _rel(g, r))
    assert _rel(g, r) < TOL
