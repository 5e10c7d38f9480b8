"""GPU tests of the overlapped-cube loop (letkf_compute, letkf_compute.F90:
156-256 with xb_add2_xa :332-382 and the cube_count division :248-255)
through the C ABI, against the FP64 oracle.

FP64 bar: 1e-10 relative on xa, axa and every cube's trans.  BF16x3
products: analysis-increment RMS error relative to the FP64 oracle's
increment, per state variable, <= 1e-5.  Several batches, both products
schedules and a hipGraph replay give the same bits as one eager batch."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from dida_letkf import engine
    engine.lib()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _inst(name):
    from dida_letkf import synth
    if name == "c0_k20_c4o2_halo":
        inst = synth.generate("c0", seed=81, nlat=18, nlon=30, n_obs=900, halo=1,
                              local_dist=900.0)
        cn, ov = 4, 2
    elif name == "c0_k20_c3o1":
        inst = synth.generate("c0", seed=82, nlat=16, nlon=25, n_obs=900, local_dist=900.0)
        cn, ov = 3, 1
    elif name == "c2_k64_c5o3":
        inst = synth.generate("c2", seed=83, nlat=20, nlon=36, nz=3, n_obs=15000,
                              local_dist=900.0)
        cn, ov = 5, 3
    elif name == "c3_k128_c4o2":
        inst = synth.generate("c3", seed=84, k=128, nlat=12, nlon=20, nz=2, n_obs=12000,
                              local_dist=900.0)
        cn, ov = 4, 2
    else:
        raise KeyError(name)
    inst.cfg.cube_n = cn
    inst.cfg.overlap = ov
    return inst


NAMES = ["c0_k20_c4o2_halo", "c0_k20_c3o1", "c2_k64_c5o3", "c3_k128_c4o2"]


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


def _run(inst, precision=None, keep=False, max_batch=0, schedule=1, graph_reps=0):
    import torch
    from dida_letkf import engine
    cfg = inst.cfg
    if precision is not None:
        cfg = copy.copy(cfg)
        cfg.precision = precision
    side = torch.cuda.Stream()
    with engine.LetkfEngine(cfg, keep_transforms=keep, max_batch_columns=max_batch) as eng, \
            torch.cuda.stream(side):
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(side.cuda_stream)
        eng.set_option(engine.OPT_SCHEDULE, schedule)
        eng.set_obs(inst.obs)
        xb, axb = _dev(inst.xb), _dev(inst.axb)
        # halo columns must come back zeroed
        xa = torch.full_like(xb, 3.0)
        axa = torch.full_like(axb, 3.0)
        outs = []
        if graph_reps:
            eng.set_option(engine.OPT_GRAPH, 1)
        for _ in range(max(1, graph_reps)):
            assert eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(),
                                      axa.data_ptr()) == 0
            side.synchronize()
            outs.append((xa.cpu().numpy(), axa.cpu().numpy(), eng.stats.graph_replay))
            xa.fill_(5.0)
            side.synchronize()
        out = dict(xa=outs[0][0], axa=outs[0][1], stats=eng.stats.as_dict(), runs=outs)
        if keep:
            out["trans"], out["w"], out["nobsl"] = eng.get_transforms()
        return out


@pytest.mark.parametrize("name", NAMES)
def test_overlap_fp64_matches_oracle(name):
    import orc
    inst = _inst(name)
    print(name, "cube_n", inst.cfg.cube_n, "overlap", inst.cfg.overlap, flush=True)
    ref = orc.letkf_compute(inst, want_trans=True)
    got = _run(inst, precision=0, keep=True)
    assert np.array_equal(got["nobsl"], ref["nobsl"])
    assert got["stats"]["n_flagged"] == 0
    e = dict(xa=rel(got["xa"], ref["xa"]), axa=rel(got["axa"], ref["axa"]),
             trans=rel(got["trans"], ref["trans"]))
    print(name, e)
    assert max(e.values()) < TOL64, e
    h = inst.cfg.halo
    if h:
        assert np.all(got["xa"][:h] == 0) and np.all(got["xa"][:, :h] == 0)
        assert np.all(got["axa"][-h:] == 0) and np.all(got["axa"][:, -h:] == 0)


def _increment_error(inst, full_m, full_r):
    cfg = inst.cfg
    h = cfg.halo
    sl = (slice(h, h + cfg.nlon), slice(h, h + cfg.nlat))
    full_b = (inst.xb + inst.axb[..., None])[sl]
    full_m, full_r = full_m[sl], full_r[sl]
    vs = cfg.var_start() - 1
    out = {}
    for v in range(cfg.n_var):
        s = slice(vs[v], vs[v] + (1 if cfg.var_lev()[v] == 1 else cfg.nz))
        num = np.sqrt(np.mean((full_m[:, :, s] - full_r[:, :, s]) ** 2))
        den = np.sqrt(np.mean((full_r[:, :, s] - full_b[:, :, s]) ** 2))
        out[v] = float(num / den)
    return out


@pytest.mark.parametrize("name", ["c2_k64_c5o3", "c3_k128_c4o2"])
def test_overlap_bf16x3_increment_error(name):
    import orc
    inst = _inst(name)
    ref = orc.letkf_compute(inst)
    got = _run(inst, precision=2)
    err = _increment_error(inst, got["xa"] + got["axa"][..., None],
                           ref["xa"] + ref["axa"][..., None])
    print(name, "BF16x3 increment err", err)
    assert max(err.values()) < 1e-5


@pytest.mark.parametrize("name", ["c0_k20_c4o2_halo", "c2_k64_c5o3"])
def test_overlap_batches_schedules_graph_same_bits(name):
    inst = _inst(name)
    one = _run(inst)
    for kw in (dict(max_batch=7), dict(schedule=0), dict(max_batch=11, schedule=0)):
        other = _run(inst, **kw)
        assert np.array_equal(one["xa"], other["xa"]), kw
        assert np.array_equal(one["axa"], other["axa"]), kw
    g = _run(inst, graph_reps=3)
    for i, (xa, axa, replay) in enumerate(g["runs"]):
        assert replay == 1
        assert np.array_equal(one["xa"], xa) and np.array_equal(one["axa"], axa), i


def test_overlap_argument_checks():
    import torch
    from dida_letkf import engine
    inst = _inst("c0_k20_c3o1")
    bad = copy.copy(inst.cfg)
    bad.overlap = bad.cube_n
    with pytest.raises(engine.DidaError):
        engine.LetkfEngine(bad)
    vert = copy.copy(inst.cfg)
    vert.vert_dist = 1.0
    with pytest.raises(engine.DidaError, match="horizontal loop"):
        engine.LetkfEngine(vert)
    cfg = copy.copy(inst.cfg)
    cfg.precision = 2
    with engine.LetkfEngine(cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.set_obs(inst.obs)
        xb = _dev(inst.xb.astype(np.float32))
        axb = _dev(inst.axb)
        xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
        with pytest.raises(engine.DidaError, match="FP64 state"):
            eng.analyze_device_f32(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr())
