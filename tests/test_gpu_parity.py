"""GPU parity: the HIP engine through the C ABI against the CPU oracle on the
same seeded inputs.  FP64 bar: 1e-10 relative (max abs error / max abs value
of the oracle output).  Reduced-precision modes: analysis-increment RMS
error relative to the FP64 oracle's increment, per state variable."""
import copy
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from dida_letkf import engine
    engine.lib()


def _run_both(inst, **kw):
    import orc
    from dida_letkf import engine
    gpu = engine.letkf_compute_nooverlap(inst, keep_transforms=True)
    ref = orc.letkf_compute(inst, want_trans=True, **kw)
    return gpu, ref


def _assert_parity(gpu, ref, tol=TOL64, what=""):
    assert gpu["n_flagged"] == 0
    assert np.array_equal(gpu["nobsl"], ref["nobsl"]), what
    e = dict(xa=rel(gpu["xa"], ref["xa"]), axa=rel(gpu["axa"], ref["axa"]),
             trans=rel(gpu["trans"], ref["trans"]),
             w=float(np.abs(gpu["overline_w"] - ref["overline_w"]).max()))
    print(what, e)
    assert e["xa"] < tol and e["axa"] < tol and e["trans"] < tol and e["w"] < tol, (what, e)


# ---------------------------------------------------------------------------
# letkf_core drop-in
# ---------------------------------------------------------------------------

def _case(k, p, seed):
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((k, p))
    Y -= Y.mean(axis=0, keepdims=True)
    rinv = (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32)
    dep = rng.standard_normal(p)
    rloc = rng.uniform(0.0, 1.0, p)
    return Y, rinv, dep, rloc


@pytest.mark.parametrize("k,p", [(2, 0), (2, 1), (3, 17), (20, 0), (20, 1), (20, 49), (32, 5),
                                 (32, 300), (64, 0), (64, 63), (64, 1000), (65, 30), (100, 2000),
                                 (128, 0), (128, 1), (128, 127), (128, 3000)])
def test_letkf_core_dropin(k, p):
    import orc
    from dida_letkf import engine
    Y, rinv, dep, rloc = _case(k, p, 100 * k + p)
    t_g, w_g = engine.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.15)
    t_o, w_o, _ = orc.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.15)
    assert rel(t_g, t_o) < TOL64
    assert np.abs(w_g - w_o).max() < TOL64 * max(1.0, np.abs(w_o).max())


def test_letkf_core_repeated_calls_reuse_context():
    """The drop-in is called once per analysis: calls that grow and shrink
    reuse the calling thread's stream and buffers and keep matching the
    oracle; an unsupported k gives NaN outputs."""
    import time
    import orc
    from dida_letkf import engine
    sizes = [(40, 300), (40, 1200), (64, 50), (40, 10), (128, 400), (20, 0), (64, 2000),
             (40, 300)] * 4
    times = []
    for i, (k, p) in enumerate(sizes):
        Y, rinv, dep, rloc = _case(k, p, 7000 + i)
        t0 = time.perf_counter()
        t_g, w_g = engine.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.1)
        times.append(time.perf_counter() - t0)
        t_o, w_o, _ = orc.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.1)
        assert rel(t_g, t_o) < TOL64, (k, p)
        assert np.abs(w_g - w_o).max() < TOL64 * max(1.0, np.abs(w_o).max())
    print("letkf_core_ per call: median %.3f ms, first %.3f ms" %
          (1e3 * float(np.median(times)), 1e3 * times[0]))
    Y, rinv, dep, rloc = _case(130, 20, 1)
    t_g, w_g = engine.letkf_core(Y, rinv, dep, rloc)
    assert np.all(np.isnan(t_g)) and np.all(np.isnan(w_g))


@pytest.mark.parametrize("k,cuts", [(100, (30, 70)), (128, (40, 90)), (128, (60, 64, 66)),
                                    (128, (8, 16, 24, 32, 40, 48, 56))])
def test_letkf_core_block_structured(k, cuts):
    """Members in groups that share no obs: A = (k-1)/rho I + Y R^-1 Y^T is
    block diagonal with the given member cuts, so the tridiagonal form has
    zero subdiagonals there and the QL log mixes chases that stay below
    column 64, cross it and stay above it (the k = 65..128 kernels' three
    chase kinds, in every order the cuts produce)."""
    import orc
    from dida_letkf import engine
    rng = np.random.default_rng(k + sum(cuts))
    edges = [0, *cuts, k]
    per = 12
    p = per * (len(edges) - 1)
    Y = np.zeros((k, p))
    for g in range(len(edges) - 1):
        Y[edges[g]:edges[g + 1], per * g:per * (g + 1)] = rng.standard_normal(
            (edges[g + 1] - edges[g], per))
    rinv = (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32)
    dep = rng.standard_normal(p)
    rloc = rng.uniform(0.0, 1.0, p)
    t_g, w_g = engine.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.15)
    t_o, w_o, _ = orc.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.15)
    assert rel(t_g, t_o) < TOL64, rel(t_g, t_o)
    assert np.abs(w_g - w_o).max() < TOL64 * max(1.0, np.abs(w_o).max())


def test_letkf_core_batched_mixed_sizes():
    import orc
    from dida_letkf import engine
    k = 40
    sizes = [0, 1, 2, 39, 40, 41, 200, 777, 3]
    cases = [_case(k, p, 7 + i) for i, p in enumerate(sizes)]
    t, w, st = engine.letkf_core_batched([c[0] for c in cases], [c[1] for c in cases],
                                         [c[2] for c in cases], [c[3] for c in cases])
    assert (st == 0).all()
    for i, c in enumerate(cases):
        t_o, w_o, _ = orc.letkf_core(*c)
        assert rel(t[i], t_o) < TOL64, i
        assert np.abs(w[i] - w_o).max() < TOL64 * max(1.0, np.abs(w_o).max()), i


# ---------------------------------------------------------------------------
# analysis loops
# ---------------------------------------------------------------------------

def test_config0_full_fp64():
    """The c0 workload (72x36, k=20, 2000 obs, GC 1000 km) in full."""
    from dida_letkf import synth
    inst = synth.generate("c0", seed=1)
    gpu, ref = _run_both(inst)
    _assert_parity(gpu, ref, what="c0")


def test_clustered_k64_reduced():
    """The c2 shape (clustered obs, 4 vars, k=64) on a reduced grid."""
    from dida_letkf import synth
    inst = synth.generate("c2", seed=3, nlat=24, nlon=48, nz=4, n_obs=20000, local_dist=900.0)
    gpu, ref = _run_both(inst)
    assert gpu["stats"]["max_rows"] > 64        # the SYRK is full rank somewhere
    _assert_parity(gpu, ref, what="c2-reduced")


def test_vertical_level_units_k32():
    """The c1 shape: vertical localisation in level units, per-level analyses."""
    from dida_letkf import synth
    inst = synth.generate("c1", seed=5, nlat=12, nlon=24, nz=6, n_obs=3000)
    assert inst.cfg.vertical
    gpu, ref = _run_both(inst)
    _assert_parity(gpu, ref, what="c1-reduced")


def test_call_statistics_match_counts():
    """The end-of-call summary (one workgroup reducing the per-analysis status
    and the per-column row counts on the device) against the per-analysis
    counts read back with the transforms and checked against the oracle's
    nobsl: row total, largest count, columns over 10,000 rows, flagged."""
    from dida_letkf import synth
    for inst in (synth.generate("c2", seed=21, k=32, nlat=18, nlon=36, nz=4, n_obs=6000),
                 synth.generate("c1", seed=22, nlat=12, nlon=24, nz=6, n_obs=3000)):
        gpu, ref = _run_both(inst)
        _assert_parity(gpu, ref, what="stats")
        st, nob = gpu["stats"], np.asarray(gpu["nobsl"], np.int64)
        nzc = nob.size // (inst.cfg.nlat * inst.cfg.nlon)
        print("nzc", nzc, "rows", st["total_rows"], "max", st["max_rows"])
        assert st["n_analyses"] == nob.size
        assert st["total_rows"] == int(nob.sum()) > 0
        assert st["max_rows"] == int(nob.max())
        assert st["rows_over_10000"] == int((nob > 10000).sum()) // nzc
        assert st["n_flagged"] == gpu["n_flagged"] == 0


def test_vertical_lnp_hybrid():
    """ln p vertical coordinate with hybrid p(ps) (ps2alnp, obs_make.F90:418)."""
    from dida_letkf import synth, config
    inst = synth.generate("c1", seed=11, nlat=10, nlon=20, nz=5, n_obs=2000)
    cfg = inst.cfg
    nz = cfg.nz
    cfg.vcoord_mode = config.VCOORD_LNP
    cfg.hyai = list(np.linspace(0.1, 0.0, nz + 1))
    cfg.hybi = list(np.linspace(0.0, 1.0, nz + 1))
    cfg.p0, cfg.ptop = 1.0e5, 200.0
    cfg.vert_dist = 0.7
    cfg.ps_index = 0
    # ps mean of ~1e5 Pa; obs ln p spread over the column
    inst.axb[..., 0] += 1.0e5
    rng = np.random.default_rng(2)
    inst.obs.vcoord = np.log(rng.uniform(1.0e4, 1.0e5, inst.obs.n_rows))
    gpu, ref = _run_both(inst)
    _assert_parity(gpu, ref, what="lnp")


def test_cube_and_halo_and_polar():
    """cube_n = 2 (one analysis per 2x2 cube), a halo of 1 (never analysed),
    polar obs dropped (letkf_yb.F90:87) and inflation != 1."""
    from dida_letkf import synth
    inst = synth.generate("c0", seed=9, nlat=18, nlon=36, n_obs=800, halo=1)
    inst.cfg.cube_n = 2
    inst.cfg.polar_discard = True
    inst.cfg.polar_lat = 60.0
    inst.cfg.inflation_factor = 1.2
    gpu, ref = _run_both(inst)
    _assert_parity(gpu, ref, what="cube-halo")
    h = inst.cfg.halo
    # halo entries untouched (zero in, zero out)
    assert np.all(gpu["xa"][:h] == 0) and np.all(gpu["axa"][:, :h] == 0)


def test_boxcar_localisation():
    from dida_letkf import synth
    inst = synth.generate("c0", seed=4, nlat=12, nlon=24, n_obs=500)
    inst.cfg.gc_filter = False
    gpu, ref = _run_both(inst)
    _assert_parity(gpu, ref, what="boxcar")


def test_no_obs_is_identity():
    """With no obs in reach, trans = sqrt(rho) I and w = 0 (p = 0 everywhere)."""
    from dida_letkf import synth, engine
    inst = synth.generate("c0", seed=2, nlat=8, nlon=16, n_obs=50)
    inst.obs.site_lat[:] = 89.9      # all obs near the pole, then dropped
    inst.cfg.polar_discard = True
    inst.cfg.polar_lat = 80.0
    gpu = engine.letkf_compute_nooverlap(inst, keep_transforms=True)
    assert gpu["stats"]["total_rows"] == 0
    assert np.abs(gpu["axa"] - inst.axb).max() < 1e-14
    assert np.abs(gpu["xa"] - inst.xb).max() < 1e-13


def test_deterministic_bitwise():
    from dida_letkf import synth, engine
    inst = synth.generate("c2", seed=6, nlat=12, nlon=24, nz=3, n_obs=3000)
    a = engine.letkf_compute_nooverlap(inst)
    b = engine.letkf_compute_nooverlap(inst)
    assert np.array_equal(a["xa"], b["xa"]) and np.array_equal(a["axa"], b["axa"])


def test_batching_does_not_change_results():
    """Column batches of any size give bitwise-equal results."""
    from dida_letkf import synth, engine
    inst = synth.generate("c0", seed=8, nlat=12, nlon=24, n_obs=700)
    outs = []
    for mb in (0, 7, 64):
        with engine.LetkfEngine(inst.cfg, max_batch_columns=mb) as eng:
            eng.set_grid(inst.grid_lat, inst.grid_lon)
            eng.set_obs(inst.obs)
            xa, axa, _ = eng.letkf_compute(inst.xb, inst.axb)
            outs.append((xa, axa))
    for xa, axa in outs[1:]:
        assert np.array_equal(xa, outs[0][0]) and np.array_equal(axa, outs[0][1])


# ---------------------------------------------------------------------------
# reduced precision (FP64 eigensolve, low-precision products)
# ---------------------------------------------------------------------------

def increment_error(inst, xa_m, axa_m, xa_r, axa_r):
    """Per-variable RMS of (member analysis error) / RMS(member increment)."""
    cfg = inst.cfg
    full_m = xa_m + axa_m[..., None]
    full_r = xa_r + axa_r[..., None]
    full_b = inst.xb + inst.axb[..., None]
    out = {}
    vs = cfg.var_start() - 1
    for v in range(cfg.n_var):
        sl = slice(vs[v], vs[v] + (1 if cfg.var_lev()[v] == 1 else cfg.nz))
        num = np.sqrt(np.mean((full_m[:, :, sl] - full_r[:, :, sl]) ** 2))
        den = np.sqrt(np.mean((full_r[:, :, sl] - full_b[:, :, sl]) ** 2))
        out[v] = float(num / den)
    return out


@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 1e-5)])
def test_reduced_precision_increment_error(prec, tol):
    import orc
    from dida_letkf import synth, engine
    inst = synth.generate("c2", seed=12, nlat=24, nlon=48, nz=4, n_obs=20000, local_dist=900.0)
    gpu = engine.letkf_compute_nooverlap(inst, precision=prec)
    ref = orc.letkf_compute(inst)
    err = increment_error(inst, gpu["xa"], gpu["axa"], ref["xa"], ref["axa"])
    print("precision", prec, "increment rel RMS err per var", err)
    assert max(err.values()) < tol


@pytest.mark.parametrize("prec", [1, 2])
def test_config3_shape_k64_reduced(prec):
    """The c3 shape at k=64 (clustered obs, sigma log-uniform in [0.3, 3]) on a
    reduced grid: FP64 at 1e-10 against the oracle, the reduced modes at the
    1e-5 increment bar."""
    import orc
    from dida_letkf import synth, engine
    inst = synth.generate("c3", seed=26, k=64, nlat=16, nlon=32, nz=3, n_obs=20000,
                          local_dist=900.0)
    ref = orc.letkf_compute(inst, want_trans=True)
    gpu64 = engine.letkf_compute_nooverlap(inst, keep_transforms=True)
    assert gpu64["stats"]["max_rows"] > 64
    _assert_parity(gpu64, ref, what="c3-k64-reduced")
    gpu = engine.letkf_compute_nooverlap(inst, precision=prec)
    err = increment_error(inst, gpu["xa"], gpu["axa"], ref["xa"], ref["axa"])
    print("c3 k=64 precision", prec, "increment rel RMS err per var", err)
    assert max(err.values()) < 1e-5


def test_epilogue_device():
    import torch
    import orc
    from dida_letkf import synth, engine
    inst = synth.generate("c0", seed=13, nlat=12, nlon=24, n_obs=400, halo=1)
    ref = orc.letkf_compute(inst)
    xo_r, sb_r, sa_r = orc.epilogue(inst, ref["xa"], ref["axa"], polar_lat=60.0, gamma1=0.6,
                                    gamma2=0.4)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    xb, axb, xa, axa = t(inst.xb), t(inst.axb), t(ref["xa"]), t(ref["axa"])
    xo = torch.zeros_like(xb)
    cfg = inst.cfg
    sb = torch.zeros((cfg.nlon, cfg.nlat, cfg.n), dtype=torch.float64, device=dev)
    sa = torch.zeros_like(sb)
    with engine.LetkfEngine(cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.epilogue_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr(), 60.0,
                            0.6, 0.4, xo.data_ptr(), sb.data_ptr(), sa.data_ptr())
        torch.cuda.synchronize()
    assert rel(xo.cpu().numpy(), xo_r) < 1e-14
    assert rel(sb.cpu().numpy(), sb_r) < 1e-14
    assert rel(sa.cpu().numpy(), sa_r) < 1e-14


@pytest.mark.parametrize("tk", [0, 1])
def test_transform_paths_match_oracle(tk):
    """Both transform implementations (split eigensolver, single kernel) on a
    k=64 clustered case and a k=20 case with p < k columns."""
    import orc
    from dida_letkf import synth, engine
    for inst in (synth.generate("c2", seed=21, nlat=16, nlon=32, nz=3, n_obs=12000,
                                local_dist=800.0),
                 synth.generate("c0", seed=22, nlat=12, nlon=24, n_obs=150)):
        gpu = engine.letkf_compute_nooverlap(inst, keep_transforms=True, transform_kernel=tk)
        ref = orc.letkf_compute(inst, want_trans=True)
        _assert_parity(gpu, ref, what=f"tk={tk} k={inst.cfg.k}")


# ---------------------------------------------------------------------------
# k = 65..128 (LDS-resident tridiagonalisation and update)
# ---------------------------------------------------------------------------

def test_config3_shape_k128_reduced():
    """The c3 shape at k=128 (clustered obs, log-uniform sigma) on a reduced grid."""
    from dida_letkf import synth
    inst = synth.generate("c3", seed=23, k=128, nlat=12, nlon=24, nz=3, n_obs=15000,
                          local_dist=900.0)
    gpu, ref = _run_both(inst)
    assert gpu["stats"]["max_rows"] > 128
    _assert_parity(gpu, ref, what="c3-k128-reduced")


def test_vertical_k100_padded():
    """k=100 (padded to 128 lanes) on the per-level vertical loop."""
    from dida_letkf import synth
    inst = synth.generate("c1", seed=24, k=100, nlat=8, nlon=16, nz=4, n_obs=2500)
    gpu, ref = _run_both(inst)
    _assert_parity(gpu, ref, what="c1-k100")


@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 1e-5)])
def test_reduced_precision_k128(prec, tol):
    import orc
    from dida_letkf import synth, engine
    inst = synth.generate("c3", seed=25, k=128, nlat=12, nlon=24, nz=3, n_obs=15000,
                          local_dist=900.0)
    gpu = engine.letkf_compute_nooverlap(inst, precision=prec)
    ref = orc.letkf_compute(inst)
    err = increment_error(inst, gpu["xa"], gpu["axa"], ref["xa"], ref["axa"])
    print("k=128 precision", prec, "increment rel RMS err per var", err)
    assert max(err.values()) < tol


@pytest.mark.parametrize("pk", [1, 2])
@pytest.mark.parametrize("prec", [0, 2])
def test_products_paths_match(pk, prec):
    """Per-analysis products (1) and tile-union products (2) against the
    oracle (FP64) or each other (BF16x3, within the mode's rounding)."""
    import orc
    from dida_letkf import synth, engine
    inst = synth.generate("c2", seed=27, nlat=20, nlon=40, nz=3, n_obs=15000, local_dist=700.0)
    gpu = engine.letkf_compute_nooverlap(inst, precision=prec, keep_transforms=True,
                                         products_kernel=pk)
    if prec == 0:
        ref = orc.letkf_compute(inst, want_trans=True)
        _assert_parity(gpu, ref, what=f"products_kernel={pk}")
    else:
        # the two paths sum the rows in different groupings: FP32-accumulation
        # rounding apart (~1e-7), each within the mode's bar of the FP64 oracle
        other = engine.letkf_compute_nooverlap(inst, precision=prec, keep_transforms=True,
                                               products_kernel=3 - pk)
        assert np.array_equal(gpu["nobsl"], other["nobsl"])
        assert rel(gpu["xa"] + gpu["axa"][..., None], other["xa"] + other["axa"][..., None]) < 1e-5
        ref = orc.letkf_compute(inst)
        err = increment_error(inst, gpu["xa"], gpu["axa"], ref["xa"], ref["axa"])
        assert max(err.values()) < 1e-4, err
