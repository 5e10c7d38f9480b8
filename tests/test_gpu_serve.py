"""Serving loop (dida_serve_*): continuous batching with HBM admission.

Each job's result is bitwise the same run alone and in a mix of concurrent
jobs; multi-cycle jobs keep the ensemble on the device and equal cycle-by-
cycle engine calls; admission refuses jobs larger than the budget; queued
jobs can be cancelled."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch, torch.device("cuda:0")


def _dev_obs(torch, dev, obs):
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).to(dev)
    d = dict(site_lat=t(obs.site_lat), site_lon=t(obs.site_lon),
             site_row_ptr=t(obs.site_row_ptr, np.int32), yb=t(obs.yb), dep=t(obs.dep),
             rinv=t(obs.rinv), vcoord=t(obs.vcoord))
    ptrs = {k: v.data_ptr() for k, v in d.items()}
    ptrs["n_sites"] = obs.n_sites
    ptrs["n_rows"] = obs.n_rows
    return d, ptrs


def _jobs(torch, dev):
    from dida_letkf import synth
    # (config, generator arguments, max_batch_columns): a k = 128 job, and a
    # job whose cycles run in several batches
    specs = [("c0", dict(seed=31, nlat=18, nlon=36, n_obs=700), 0),
             ("c2", dict(seed=32, nlat=16, nlon=32, nz=3, n_obs=9000, local_dist=800.0), 0),
             ("c1", dict(seed=33, nlat=10, nlon=20, nz=4, n_obs=2500), 0),
             ("c0", dict(seed=34, nlat=12, nlon=24, n_obs=400), 0),
             ("c3", dict(seed=35, k=128, nlat=10, nlon=20, nz=2, n_obs=8000, local_dist=900.0), 0),
             ("c2", dict(seed=36, nlat=20, nlon=40, nz=2, n_obs=6000, local_dist=700.0), 96)]
    jobs = []
    for name, kw, mb in specs:
        a = synth.generate(name, **kw)
        kw2 = dict(kw)
        kw2["seed"] = kw["seed"] + 100          # second cycle's obs
        b = synth.generate(name, **kw2)
        t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        o0, p0 = _dev_obs(torch, dev, a.obs)
        o1, p1 = _dev_obs(torch, dev, b.obs)
        jobs.append(dict(inst=a, xb=t(a.xb), axb=t(a.axb), obs=[p0, p1], keep=[o0, o1], mb=mb))
    return jobs


def _run(server_cls, torch, jobs, max_active, budget=64 << 30):
    outs = []
    with server_cls(budget, max_active) as srv:
        ids = []
        for j in jobs:
            xa = torch.zeros_like(j["xb"])
            axa = torch.zeros_like(j["axb"])
            jid = srv.submit(j["inst"].cfg, j["inst"].grid_lat, j["inst"].grid_lon,
                             j["xb"].data_ptr(), j["axb"].data_ptr(), j["obs"], xa.data_ptr(),
                             axa.data_ptr(), max_batch_columns=j.get("mb", 0))
            ids.append((jid, xa, axa))
            if max_active == 1:
                st = srv.wait(jid)
                assert st["state"] == 2, st
        for jid, xa, axa in ids:
            st = srv.wait(jid)
            assert st["state"] == 2 and st["rc"] == 0, st
            outs.append((xa.cpu().numpy(), axa.cpu().numpy(), st))
    return outs


def test_alone_equals_mixed(torch_dev):
    torch, dev = torch_dev
    from dida_letkf import engine
    from dida_letkf.engine import Server
    jobs = _jobs(torch, dev)
    alone = _run(Server, torch, jobs, max_active=1)
    mixed = _run(Server, torch, jobs, max_active=4)
    for (xa1, axa1, s1), (xa2, axa2, s2) in zip(alone, mixed):
        assert np.array_equal(xa1, xa2) and np.array_equal(axa1, axa2)
    # the same jobs through an engine handle, cycle by cycle: same bits
    for j, (xa1, axa1, _) in zip(jobs, alone):
        with engine.LetkfEngine(j["inst"].cfg, max_batch_columns=j["mb"]) as eng:
            eng.set_grid(j["inst"].grid_lat, j["inst"].grid_lon)
            eng.set_stream(torch.cuda.current_stream().cuda_stream)
            bx, ba = j["xb"], j["axb"]
            for o in j["obs"]:
                ox, oa = torch.zeros_like(bx), torch.zeros_like(ba)
                eng.set_obs_device(o["n_sites"], o["site_lat"], o["site_lon"],
                                   o["site_row_ptr"], o["n_rows"], o["yb"], o["dep"], o["rinv"],
                                   o["vcoord"])
                eng.analyze_device(bx.data_ptr(), ba.data_ptr(), ox.data_ptr(), oa.data_ptr())
                bx, ba = ox, oa
            torch.cuda.synchronize()
        assert np.array_equal(bx.cpu().numpy(), xa1) and np.array_equal(ba.cpu().numpy(), axa1)
    # jobs ran concurrently in the mixed run: some pair of [start, end] overlaps
    iv = [(s[2]["t_start"], s[2]["t_end"]) for s in mixed]
    assert any(a[0] < b[1] and b[0] < a[1] for i, a in enumerate(iv) for b in iv[i + 1:])


def test_cycles_equal_engine_calls(torch_dev):
    torch, dev = torch_dev
    from dida_letkf import synth, engine
    inst = synth.generate("c0", seed=41, nlat=12, nlon=24, n_obs=500)
    inst2 = synth.generate("c0", seed=42, nlat=12, nlon=24, n_obs=500)
    # engine, cycle by cycle, on the host
    with engine.LetkfEngine(inst.cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_obs(inst.obs)
        xa0, axa0, _ = eng.letkf_compute(inst.xb, inst.axb)
        eng.set_obs(inst2.obs)
        xa1, axa1, _ = eng.letkf_compute(xa0, axa0)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    o0, p0 = _dev_obs(torch, dev, inst.obs)
    o1, p1 = _dev_obs(torch, dev, inst2.obs)
    xb, axb = t(inst.xb), t(inst.axb)
    xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
    with engine.Server(8 << 30, 2) as srv:
        jid = srv.submit(inst.cfg, inst.grid_lat, inst.grid_lon, xb.data_ptr(), axb.data_ptr(),
                         [p0, p1], xa.data_ptr(), axa.data_ptr())
        st = srv.wait(jid)
    assert st["state"] == 2 and st["n_analyses"] == 2 * inst.cfg.n_analyses()
    assert np.array_equal(xa.cpu().numpy(), xa1) and np.array_equal(axa.cpu().numpy(), axa1)


def test_admission_refusal_and_cancel(torch_dev):
    torch, dev = torch_dev
    from dida_letkf import synth, engine
    inst = synth.generate("c0", seed=51, nlat=12, nlon=24, n_obs=300)
    o0, p0 = _dev_obs(torch, dev, inst.obs)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    xb, axb = t(inst.xb), t(inst.axb)
    need = engine.job_bytes(inst.cfg, inst.obs.n_sites, inst.obs.n_rows)
    assert need > 0
    with engine.Server(need // 2, 1) as srv:
        with pytest.raises(engine.DidaError) as ei:
            srv.submit(inst.cfg, inst.grid_lat, inst.grid_lon, xb.data_ptr(), axb.data_ptr(),
                       [p0], torch.zeros_like(xb).data_ptr(), torch.zeros_like(axb).data_ptr())
        assert "(-5)" in str(ei.value)
    outs = [(torch.zeros_like(xb), torch.zeros_like(axb)) for _ in range(4)]
    # room for exactly one job at a time (regions are whole 256-byte granules)
    with engine.Server(need + 512, 1) as srv:
        ids = [srv.submit(inst.cfg, inst.grid_lat, inst.grid_lon, xb.data_ptr(), axb.data_ptr(),
                          [p0, p0, p0], xa.data_ptr(), axa.data_ptr()) for xa, axa in outs]
        srv.cancel(ids[-1])
        sts = [srv.wait(i) for i in ids]
    assert [s["state"] for s in sts[:3]] == [2, 2, 2]
    assert sts[3]["state"] == 4
    # FIFO with room for one: job i starts after job i-1 ended
    for a, b in zip(sts[:2], sts[1:3]):
        assert b["t_start"] >= a["t_end"]


def test_operator_mode_cycles_equal_engine_calls(torch_dev):
    """Serving jobs whose cycles carry raw obs (row_var, row_ph, yo): the
    server runs the observation operator on each cycle's device-resident
    background; the result equals operator + analysis called cycle by cycle."""
    torch, dev = torch_dev
    from dida_letkf import synth, engine
    c = synth.obs_operator_case("c2", seed=61, k=24, nlat=12, nlon=24, nz=4, n_obs=1500,
                                n_stations=10, n_ps=30)
    inst = c.inst
    cfg = inst.cfg
    cfg.local_dist = 900.0
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).to(dev)
    nr = int(c.site_row_ptr[-1])
    keep = dict(la=t(c.site_lat), lo=t(c.site_lon), rp=t(c.site_row_ptr, np.int32),
                rv=t(c.row_var, np.int32), rph=t(c.row_ph), rinv=t(np.full(nr, 0.5)),
                vco=t(np.log(c.row_ph)))
    yos = [t(c.yo + 0.3 * i) for i in range(3)]
    cyc = [dict(n_sites=c.site_lat.size, site_lat=keep["la"].data_ptr(),
                site_lon=keep["lo"].data_ptr(), site_row_ptr=keep["rp"].data_ptr(), n_rows=nr,
                rinv=keep["rinv"].data_ptr(), vcoord=keep["vco"].data_ptr(),
                row_var=keep["rv"].data_ptr(), row_ph=keep["rph"].data_ptr(), yo=y.data_ptr())
           for y in yos]
    xb, axb = t(inst.xb), t(inst.axb)
    # engine, cycle by cycle
    yb = torch.empty((nr, cfg.k), dtype=torch.float64, device=dev)
    dep = torch.empty(nr, dtype=torch.float64, device=dev)
    bx, ba = xb, axb
    with engine.LetkfEngine(cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        for o in cyc:
            ox, oa = torch.zeros_like(xb), torch.zeros_like(axb)
            eng.obs_operator_device(bx.data_ptr(), ba.data_ptr(), o["n_sites"], o["site_lat"],
                                    o["site_lon"], o["site_row_ptr"], nr, o["row_var"],
                                    o["row_ph"], o["yo"], yb.data_ptr(), dep.data_ptr())
            eng.set_obs_device(o["n_sites"], o["site_lat"], o["site_lon"], o["site_row_ptr"], nr,
                               yb.data_ptr(), dep.data_ptr(), o["rinv"], o["vcoord"])
            eng.analyze_device(bx.data_ptr(), ba.data_ptr(), ox.data_ptr(), oa.data_ptr())
            eng.fill_halo_device(ox.data_ptr(), oa.data_ptr())
            bx, ba = ox, oa
        torch.cuda.synchronize()
    ref_x, ref_a = bx.cpu().numpy(), ba.cpu().numpy()
    xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
    with engine.Server(16 << 30, 2) as srv:
        jid = srv.submit(cfg, inst.grid_lat, inst.grid_lon, xb.data_ptr(), axb.data_ptr(), cyc,
                         xa.data_ptr(), axa.data_ptr())
        st = srv.wait(jid)
    assert st["state"] == 2 and st["rc"] == 0, st
    # the last cycle's halo is filled too
    assert np.array_equal(xa.cpu().numpy(), ref_x) and np.array_equal(axa.cpu().numpy(), ref_a)
