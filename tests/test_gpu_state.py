"""GPU tests of the FP32 perturbation state (f32 MFMA update) and of the
caller epilogue fused into the update (module_da_parallel.F90:219-261),
through the C ABI, against the FP64 oracle.

FP64 bar: 1e-10 relative.  FP32 state: analysis-increment RMS error relative
to the FP64 oracle's increment, per state variable, <= 1e-5 (the inputs
themselves are the FP64 background rounded to FP32)."""
import os

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


def _dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a if dtype is None else a.astype(dtype))
    return torch.from_numpy(a).to(torch.device("cuda:0"))


def _engine(inst, precision=None):
    import copy
    import torch
    from dida_letkf import engine
    cfg = inst.cfg
    if precision is not None:
        cfg = copy.copy(cfg)
        cfg.precision = precision
    eng = engine.LetkfEngine(cfg)
    eng.set_grid(inst.grid_lat, inst.grid_lon)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    eng.set_obs(inst.obs)
    return eng


def increment_error(inst, full_m, full_r):
    """Per-variable RMS of (analysis difference) / RMS(FP64 increment), over
    interior columns."""
    cfg = inst.cfg
    h = cfg.halo
    sl_c = (slice(h, h + cfg.nlon), slice(h, h + cfg.nlat))
    full_b = (inst.xb + inst.axb[..., None])[sl_c]
    full_m, full_r = full_m[sl_c], full_r[sl_c]
    out = {}
    vs = cfg.var_start() - 1
    for v in range(cfg.n_var):
        sl = slice(vs[v], vs[v] + (1 if cfg.var_lev()[v] == 1 else cfg.nz))
        num = np.sqrt(np.mean((full_m[:, :, sl] - full_r[:, :, sl]) ** 2))
        den = np.sqrt(np.mean((full_r[:, :, sl] - full_b[:, :, sl]) ** 2))
        out[v] = float(num / den)
    return out


def _run_f32(inst, precision):
    import torch
    with _engine(inst, precision) as eng:
        xb = _dev(inst.xb, np.float32)
        axb = _dev(inst.axb)
        xa = torch.zeros_like(xb)
        axa = torch.zeros_like(axb)
        rc = eng.analyze_device_f32(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr())
        torch.cuda.synchronize()
        return rc, xa.cpu().numpy().astype(np.float64), axa.cpu().numpy()


def _instances():
    from dida_letkf import synth
    return {
        "c2_k64": synth.generate("c2", seed=31, nlat=24, nlon=48, nz=4, n_obs=20000,
                                 local_dist=900.0),
        "c1_k32_vertical": synth.generate("c1", seed=32, nlat=12, nlon=24, nz=6, n_obs=3000),
        "c3_k128": synth.generate("c3", seed=33, k=128, nlat=10, nlon=20, nz=3, n_obs=12000,
                                  local_dist=900.0),
        "c0_k20_halo": synth.generate("c0", seed=34, nlat=18, nlon=36, n_obs=800, halo=1),
    }


@pytest.mark.parametrize("name", ["c2_k64", "c1_k32_vertical", "c3_k128", "c0_k20_halo"])
@pytest.mark.parametrize("prec", [0, 2])
def test_f32_state_increment_error(name, prec):
    import orc
    inst = _instances()[name]
    ref = orc.letkf_compute(inst)
    rc, xa, axa = _run_f32(inst, prec)
    assert rc == 0
    err = increment_error(inst, xa + axa[..., None], ref["xa"] + ref["axa"][..., None])
    print(name, "precision", prec, "FP32 state increment rel RMS err per var", err)
    assert max(err.values()) < 1e-5
    # the perturbations stay centred (the means are FP64)
    h = inst.cfg.halo
    inner = (slice(h, h + inst.cfg.nlon), slice(h, h + inst.cfg.nlat))
    assert np.abs(xa[inner].mean(axis=-1)).max() < 1e-5 * np.abs(xa[inner]).max()


def _epilogue_device(inst, polar_lat, g1, g2, f32=False, precision=None):
    import torch
    cfg = inst.cfg
    with _engine(inst, precision) as eng:
        xb = _dev(inst.xb, np.float32 if f32 else None)
        axb = _dev(inst.axb)
        xo = torch.full(inst.xb.shape, 7.0, dtype=torch.float64, device=axb.device)
        sb = torch.zeros((cfg.nlon, cfg.nlat, cfg.n), dtype=torch.float64, device=axb.device)
        sa = torch.zeros_like(sb)
        rc = eng.analyze_epilogue_device(xb.data_ptr(), axb.data_ptr(), polar_lat, g1, g2,
                                         xo.data_ptr(), sb.data_ptr(), sa.data_ptr(), f32=f32)
        torch.cuda.synchronize()
        return rc, xo.cpu().numpy(), sb.cpu().numpy(), sa.cpu().numpy()


@pytest.mark.parametrize("name", ["c2_k64", "c1_k32_vertical", "c3_k128", "c0_k20_halo"])
def test_fused_epilogue_fp64(name):
    import orc
    inst = _instances()[name]
    ref = orc.letkf_compute(inst)
    for pl, g1, g2 in ((90.0, 0.0, 1.0), (45.0, 0.3, 0.7)):
        xo_r, sb_r, sa_r = orc.epilogue(inst, ref["xa"], ref["axa"], polar_lat=pl, gamma1=g1,
                                        gamma2=g2)
        rc, xo, sb, sa = _epilogue_device(inst, pl, g1, g2)
        assert rc == 0
        e = dict(xa_out=rel(xo, xo_r), spread_b=rel(sb, sb_r), spread_a=rel(sa, sa_r))
        print(name, pl, g1, g2, e)
        assert max(e.values()) < TOL64, e
        h = inst.cfg.halo
        if h:
            assert np.all(xo[:h] == 0) and np.all(xo[:, :h] == 0)


@pytest.mark.parametrize("name", ["c2_k64", "c3_k128"])
def test_fused_epilogue_f32_state(name):
    import orc
    inst = _instances()[name]
    ref = orc.letkf_compute(inst)
    xo_r, sb_r, sa_r = orc.epilogue(inst, ref["xa"], ref["axa"], polar_lat=60.0, gamma1=0.2,
                                    gamma2=0.8)
    rc, xo, sb, sa = _epilogue_device(inst, 60.0, 0.2, 0.8, f32=True, precision=2)
    assert rc == 0
    err = increment_error(inst, xo, xo_r)
    print(name, "FP32 state + BF16x3 products, fused epilogue, increment err", err)
    assert max(err.values()) < 1e-5
    assert rel(sb, sb_r) < 1e-6 and rel(sa, sa_r) < 1e-5


def test_fused_epilogue_failed_analyses_keep_background():
    """Analyses that fail both eigensolver passes give the background and equal
    spreads in the fused epilogue's outputs."""
    from dida_letkf import synth
    inst = synth.generate("c0", seed=35, nlat=6, nlon=12, n_obs=1500, local_dist=900.0)
    old = {v: os.environ.get(v) for v in ("DIDA_ROT_CAP", "DIDA_RERUN_ROT_CAP")}
    os.environ["DIDA_ROT_CAP"] = "16"
    os.environ["DIDA_RERUN_ROT_CAP"] = "16"
    try:
        rc, xo, sb, sa = _epilogue_device(inst, 90.0, 0.0, 1.0)
    finally:
        for v, x in old.items():
            if x is None:
                os.environ.pop(v, None)
            else:
                os.environ[v] = x
    assert rc > 0
    full_b = inst.xb + inst.axb[..., None]
    same = np.all(xo == full_b, axis=(2, 3))
    assert same.sum() == rc                       # one column per failed analysis
    assert np.array_equal(sa[same], sb[same])


# ---------------------------------------------------------------------------
# the caller's split on the device (module_da_parallel.F90:186-211) and the
# whole run_da_parallel analysis step (:186-261)
# ---------------------------------------------------------------------------

def _split_np(x_in):
    k = x_in.shape[-1]
    axb = x_in.sum(axis=-1) / k
    return x_in - axb[..., None], axb


def test_split_device_fp64_and_f32():
    import torch
    inst = _instances()["c0_k20_halo"]
    x_in = inst.xb + inst.axb[..., None] + 3.0
    xb_r, axb_r = _split_np(x_in)
    with _engine(inst) as eng:
        xi = _dev(x_in)
        xb = torch.zeros_like(xi)
        xb32 = torch.zeros(xi.shape, dtype=torch.float32, device=xi.device)
        axb = torch.zeros(x_in.shape[:-1], dtype=torch.float64, device=xi.device)
        assert eng.split_device(xi.data_ptr(), xb_ptr=xb.data_ptr(), axb_ptr=axb.data_ptr()) == 0
        torch.cuda.synchronize()
        assert rel(axb.cpu().numpy(), axb_r) < 1e-14
        assert np.abs(xb.cpu().numpy() - xb_r).max() < 1e-13 * np.abs(x_in).max()
        # the perturbations are centred
        assert np.abs(xb.cpu().numpy().sum(axis=-1)).max() < 1e-12 * np.abs(x_in).max()
        axb.zero_()
        assert eng.split_device(xi.data_ptr(), xb32_ptr=xb32.data_ptr(),
                                axb_ptr=axb.data_ptr()) == 0
        torch.cuda.synchronize()
        assert rel(axb.cpu().numpy(), axb_r) < 1e-14
        x32 = xb32.cpu().numpy().astype(np.float64)
        assert np.abs(x32 - xb_r).max() <= 2 ** -23 * np.abs(xb_r).max() * 1.01


def test_split_device_nan_is_an_error():
    import torch
    from dida_letkf import engine
    inst = _instances()["c0_k20_halo"]
    h = inst.cfg.halo
    x_in = inst.xb + inst.axb[..., None]
    # C order (lon, lat, n, k): Fortran xb_in(e, k, j, i) with j = lat, i = lon
    x_in[5, 0, 0, 3] = np.nan
    x_in[9, 7, 0, 0] = np.nan
    with _engine(inst) as eng:
        xi = _dev(x_in)
        xb = torch.zeros_like(xi)
        axb = torch.zeros(x_in.shape[:-1], dtype=torch.float64, device=xi.device)
        with pytest.raises(engine.DidaError) as ei:
            eng.split_device(xi.data_ptr(), xb_ptr=xb.data_ptr(), axb_ptr=axb.data_ptr())
    msg = str(ei.value)
    print(msg)
    assert "in 2 state elements" in msg
    assert f"k=1 j={0 + 1 - h} i={5 + 1 - h}" in msg


@pytest.mark.parametrize("name", ["c2_k64", "c0_k20_halo", "c3_k128"])
def test_run_da_device_fp64(name):
    import copy
    import orc
    import torch
    inst = _instances()[name]
    x_in = inst.xb + inst.axb[..., None]
    xb_r, axb_r = _split_np(x_in)
    ref_inst = copy.copy(inst)
    ref_inst.xb, ref_inst.axb = xb_r, axb_r
    ref = orc.letkf_compute(ref_inst)
    xo_r, sb_r, sa_r = orc.epilogue(ref_inst, ref["xa"], ref["axa"], polar_lat=60.0, gamma1=0.25,
                                    gamma2=0.75)
    cfg = inst.cfg
    with _engine(inst) as eng:
        xi = _dev(x_in)
        xo = torch.full(x_in.shape, 7.0, dtype=torch.float64, device=xi.device)
        sb = torch.zeros((cfg.nlon, cfg.nlat, cfg.n), dtype=torch.float64, device=xi.device)
        sa = torch.zeros_like(sb)
        assert eng.run_da_device(xi.data_ptr(), 60.0, 0.25, 0.75, xo.data_ptr(), sb.data_ptr(),
                                 sa.data_ptr()) == 0
        torch.cuda.synchronize()
        e = dict(xa_out=rel(xo.cpu().numpy(), xo_r), spread_b=rel(sb.cpu().numpy(), sb_r),
                 spread_a=rel(sa.cpu().numpy(), sa_r))
    print(name, e)
    assert max(e.values()) < TOL64, e


@pytest.mark.parametrize("name", ["c2_k64", "c3_k128"])
def test_run_da_device_f32_state(name):
    import copy
    import orc
    import torch
    inst = _instances()[name]
    x_in = inst.xb + inst.axb[..., None]
    xb_r, axb_r = _split_np(x_in)
    ref_inst = copy.copy(inst)
    ref_inst.xb, ref_inst.axb = xb_r, axb_r
    ref = orc.letkf_compute(ref_inst)
    xo_r, sb_r, sa_r = orc.epilogue(ref_inst, ref["xa"], ref["axa"], polar_lat=60.0, gamma1=0.2,
                                    gamma2=0.8)
    cfg = inst.cfg
    with _engine(inst, precision=2) as eng:
        xi = _dev(x_in)
        xo = torch.zeros(x_in.shape, dtype=torch.float64, device=xi.device)
        sb = torch.zeros((cfg.nlon, cfg.nlat, cfg.n), dtype=torch.float64, device=xi.device)
        sa = torch.zeros_like(sb)
        assert eng.run_da_device(xi.data_ptr(), 60.0, 0.2, 0.8, xo.data_ptr(), sb.data_ptr(),
                                 sa.data_ptr(), state_f32=True) == 0
        torch.cuda.synchronize()
        xo, sb, sa = xo.cpu().numpy(), sb.cpu().numpy(), sa.cpu().numpy()
    err = increment_error(ref_inst, xo, xo_r)
    print(name, "run_da FP32 state + BF16x3 increment err", err)
    assert max(err.values()) < 1e-5
    assert rel(sb, sb_r) < 1e-6 and rel(sa, sa_r) < 1e-5


def test_run_da_device_nan_fails():
    import torch
    from dida_letkf import engine
    inst = _instances()["c0_k20_halo"]
    cfg = inst.cfg
    x_in = inst.xb + inst.axb[..., None]
    x_in[3, 4, 0, 2] = np.inf - np.inf
    with _engine(inst) as eng:
        xi = _dev(x_in)
        xo = torch.zeros(x_in.shape, dtype=torch.float64, device=xi.device)
        sb = torch.zeros((cfg.nlon, cfg.nlat, cfg.n), dtype=torch.float64, device=xi.device)
        sa = torch.zeros_like(sb)
        with pytest.raises(engine.DidaError, match="background NaN in 1 state elements"):
            eng.run_da_device(xi.data_ptr(), 90.0, 0.0, 1.0, xo.data_ptr(), sb.data_ptr(),
                              sa.data_ptr())
