"""GPU tests of the products schedule and the capture-ready call, through the
C ABI:

* length buckets (in dispatch order, or on persistent work queues) against
  the column-order launch, bit for bit (FP64, FP32, BF16 at k <= 64 with one
  and two row groups per step, BF16 at k = 128, the vertical loop, several
  batches);
* the bucket statistics against the local obs counts (same edges as
  ``engine.BUCKET_EDGES``) and the in-kernel timing's sanity;
* a row-list capacity that is too small: the call runs again with the
  measured need and gives the same bits;
* hipGraph capture: several replays, with eager device work and
  synchronisations between them, each equal to the eager call bit for bit,
  and a new capture when a buffer changes;
* the tuner: writes its cache and leaves the results unchanged.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from dida_letkf import engine
    engine.lib()


def _instances():
    from dida_letkf import synth
    return {
        "c2_k64": synth.generate("c2", seed=41, nlat=24, nlon=48, nz=4, n_obs=20000,
                                 local_dist=900.0),
        "c1_k32_vertical": synth.generate("c1", seed=42, nlat=12, nlon=24, nz=6, n_obs=3000),
        "c3_k128": synth.generate("c3", seed=43, k=128, nlat=10, nlon=20, nz=3, n_obs=12000,
                                  local_dist=900.0),
        # >= 2048 analyses: the eigensolver's half-batch pipeline on two streams
        "c0_k64_pipe": synth.generate("c0", seed=44, k=64, nlat=48, nlon=48, n_obs=400,
                                      local_dist=600.0),
        "c0_k128_pipe": synth.generate("c0", seed=45, k=128, nlat=48, nlon=48, n_obs=400,
                                       local_dist=600.0),
    }


def _dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a if dtype is None else a.astype(dtype))
    return torch.from_numpy(a).to(torch.device("cuda:0"))


class _Env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        for k, v in self.kv.items():
            os.environ[k] = str(v)

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _run(inst, precision=None, schedule=1, variant=None, f32=False, row_cap=0, max_batch=0,
         timing=0, keep=False):
    """One analysis on a fresh handle; returns (xa, axa, stats dict, nobsl)."""
    import copy
    import torch
    from dida_letkf import engine
    cfg = inst.cfg
    if precision is not None:
        cfg = copy.copy(cfg)
        cfg.precision = precision
    env = {} if variant is None else {"DIDA_PRODUCTS_VARIANT": variant}
    with _Env(**env):
        eng = engine.LetkfEngine(cfg, keep_transforms=keep, max_batch_columns=max_batch)
    with eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.set_option(engine.OPT_SCHEDULE, schedule)
        eng.set_option(engine.OPT_TIMING, timing)
        if row_cap:
            eng.set_option(engine.OPT_ROW_CAP, row_cap)
        eng.set_obs(inst.obs)
        xb = _dev(inst.xb, np.float32 if f32 else None)
        axb = _dev(inst.axb)
        xa = torch.zeros_like(xb)
        axa = torch.zeros_like(axb)
        fn = eng.analyze_device_f32 if f32 else eng.analyze_device
        rc = fn(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr())
        torch.cuda.synchronize()
        assert rc == 0
        nob = eng.get_transforms()[2] if keep else None
        return xa.cpu().numpy(), axa.cpu().numpy(), eng.stats.as_dict(), nob


# variants: 1, 2 = dispatch order with 1 or 2 row groups per step, 3, 4 = the
# same on persistent work queues (two row groups: BF16 at k <= 64 only)
CASES = [("c2_k64", 2, None, False), ("c2_k64", 2, 2, False), ("c2_k64", 2, 3, False),
         ("c2_k64", 2, 4, False), ("c2_k64", 0, None, False), ("c2_k64", 0, 3, False),
         ("c2_k64", 1, 3, False), ("c2_k64", 2, 2, True), ("c1_k32_vertical", 2, 2, False),
         ("c1_k32_vertical", 0, 3, False), ("c3_k128", 2, None, False), ("c3_k128", 2, 3, False),
         ("c3_k128", 0, 3, False)]


@pytest.mark.parametrize("name,prec,variant,f32", CASES)
def test_queue_equals_column_order(name, prec, variant, f32):
    inst = _instances()[name]
    print(name, "precision", prec, "variant", variant, "f32 state", f32, flush=True)
    xa0, axa0, s0, _ = _run(inst, prec, schedule=0, f32=f32)
    xa1, axa1, s1, _ = _run(inst, prec, schedule=1, variant=variant, f32=f32)
    assert s0["schedule"] == 0 and s1["schedule"] == 1
    if variant:
        valid = (1, 2, 3, 4) if inst.cfg.k <= 64 and prec == 2 else (1, 3)
        used = [v for v, n in zip(s1["variant"], s1["bucket_analyses"]) if n]
        assert set(used) == ({variant} if variant in valid else {1})
    assert np.array_equal(xa0, xa1) and np.array_equal(axa0, axa1)
    assert s0["total_rows"] == s1["total_rows"]


def test_queue_several_batches_and_bucket_stats():
    from dida_letkf import engine
    inst = _instances()["c2_k64"]
    xa0, axa0, s0, nob = _run(inst, 2, schedule=0, keep=True)
    xa1, axa1, s1, nob1 = _run(inst, 2, schedule=1, max_batch=200, timing=1, keep=True)
    xa2, axa2, s2, _ = _run(inst, 2, schedule=1, max_batch=200, timing=1, variant=3)
    assert np.array_equal(xa0, xa2) and np.array_equal(axa0, axa2)
    assert 0.0 < s2["products_balance"] <= 1.0 + 1e-9
    assert np.array_equal(xa0, xa1) and np.array_equal(axa0, axa1)
    assert np.array_equal(nob, nob1)
    edges = np.array(engine.BUCKET_EDGES)
    want_n = np.zeros(8, np.int64)
    want_r = np.zeros(8, np.int64)
    for p in nob:
        b = 0 if p == 0 else min(7, int(np.searchsorted(edges, p, side="left")))
        want_n[b] += 1
        want_r[b] += p
    print("buckets", want_n, "rows", want_r, "wave ms", s1["bucket_wave_ms"],
          "balance", s1["products_balance"])
    assert np.array_equal(want_n, s1["bucket_analyses"])
    assert np.array_equal(want_r, s1["bucket_rows"])
    assert sum(s1["bucket_analyses"]) == s1["n_analyses"]
    assert sum(s1["bucket_rows"]) == s1["total_rows"]
    assert 0.0 < s1["products_balance"] <= 1.0 + 1e-9
    ms = np.array(s1["bucket_wave_ms"])
    assert np.all(ms >= 0) and ms.sum() > 0
    assert np.all(ms[np.array(s1["bucket_analyses"]) == 0] == 0)


@pytest.mark.parametrize("schedule,max_batch", [(1, 0), (0, 0), (1, 200)])
def test_row_capacity_overflow_reruns(schedule, max_batch):
    inst = _instances()["c2_k64"]
    xa0, axa0, s0, _ = _run(inst, 2, schedule=schedule, max_batch=max_batch)
    assert s0["reran_rows"] == 0
    xa1, axa1, s1, _ = _run(inst, 2, schedule=schedule, max_batch=max_batch, row_cap=1000)
    print("row_cap", s0["row_cap"], "->", s1["row_cap"], "reran", s1["reran_rows"])
    assert s1["reran_rows"] == 1 and s1["row_cap"] > 1000
    assert np.array_equal(xa0, xa1) and np.array_equal(axa0, axa1)


@pytest.mark.parametrize("name,prec,f32", [("c2_k64", 2, True), ("c2_k64", 0, False),
                                           ("c1_k32_vertical", 2, False),
                                           ("c3_k128", 2, False), ("c0_k64_pipe", 2, True),
                                           ("c0_k128_pipe", 0, False)])
def test_graph_replays_equal_eager(name, prec, f32):
    import copy
    import torch
    from dida_letkf import engine
    inst = _instances()[name]
    print(name, "precision", prec, "f32", f32, flush=True)
    xa_e, axa_e, _, _ = _run(inst, prec, f32=f32)
    cfg = copy.copy(inst.cfg)
    cfg.precision = prec
    side = torch.cuda.Stream()
    with engine.LetkfEngine(cfg) as eng, torch.cuda.stream(side):
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(side.cuda_stream)
        eng.set_option(engine.OPT_GRAPH, 1)
        eng.set_obs(inst.obs)
        xb = _dev(inst.xb, np.float32 if f32 else None)
        axb = _dev(inst.axb)
        xa = torch.zeros_like(xb)
        axa = torch.zeros_like(axb)
        fn = eng.analyze_device_f32 if f32 else eng.analyze_device
        junk = torch.randn(1 << 20, device="cuda:0")
        for rep in range(4):
            xa.fill_(float("nan"))
            axa.fill_(float("nan"))
            side.synchronize()
            assert fn(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr()) == 0
            side.synchronize()
            st = eng.stats.as_dict()
            assert st["graph_replay"] == 1 and st["reran_rows"] == 0
            assert np.array_equal(xa.cpu().numpy(), xa_e), rep
            assert np.array_equal(axa.cpu().numpy(), axa_e), rep
            # eager device work between the replays
            junk = (junk * 1.0001 + 1.0).sin_()
            torch.cuda.synchronize()
        # a new output buffer: captured again, same result
        xa2 = torch.zeros_like(xb)
        assert fn(xb.data_ptr(), axb.data_ptr(), xa2.data_ptr(), axa.data_ptr()) == 0
        side.synchronize()
        assert np.array_equal(xa2.cpu().numpy(), xa_e)


def test_graph_after_capacity_rerun():
    """A graph call whose lists overflow runs again eagerly; the next call
    captures with the grown capacity."""
    import torch
    from dida_letkf import engine
    inst = _instances()["c2_k64"]
    xa_e, axa_e, _, _ = _run(inst)
    side = torch.cuda.Stream()
    with engine.LetkfEngine(inst.cfg) as eng, torch.cuda.stream(side):
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(side.cuda_stream)
        eng.set_option(engine.OPT_GRAPH, 1)
        eng.set_option(engine.OPT_ROW_CAP, 500)
        eng.set_obs(inst.obs)
        xb, axb = _dev(inst.xb), _dev(inst.axb)
        xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
        assert eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr()) == 0
        side.synchronize()
        assert eng.stats.reran_rows == 1 and eng.stats.graph_replay == 0
        assert np.array_equal(xa.cpu().numpy(), xa_e)
        for _ in range(2):
            xa.zero_()
            side.synchronize()
            assert eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(),
                                      axa.data_ptr()) == 0
            side.synchronize()
            assert eng.stats.graph_replay == 1 and eng.stats.reran_rows == 0
            assert np.array_equal(xa.cpu().numpy(), xa_e)


def test_tuner_writes_cache_and_keeps_results(tmp_path):
    import torch
    from dida_letkf import engine
    import copy
    inst = _instances()["c2_k64"]
    xa_e, axa_e, _, _ = _run(inst, 2)
    cfg = copy.copy(inst.cfg)
    cfg.precision = 2
    cache = tmp_path / "tune.tsv"
    with _Env(DIDA_TUNE_FILE=str(cache)):
        with engine.LetkfEngine(cfg) as eng:
            eng.set_grid(inst.grid_lat, inst.grid_lon)
            eng.set_stream(torch.cuda.current_stream().cuda_stream)
            eng.set_option(engine.OPT_TUNE, 1)
            eng.set_obs(inst.obs)
            xb, axb = _dev(inst.xb), _dev(inst.axb)
            xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
            assert eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(),
                                      axa.data_ptr()) == 0
            torch.cuda.synchronize()
            st = eng.stats.as_dict()
    assert np.array_equal(xa.cpu().numpy(), xa_e) and np.array_equal(axa.cpu().numpy(), axa_e)
    lines = cache.read_text().split("\n")
    rows = [ln.split() for ln in lines if ln.strip()]
    print(rows)
    nonempty = [b for b in range(8) if st["bucket_analyses"][b]]
    got = {int(r[3]): int(r[4]) for r in rows if int(r[1]) == inst.cfg.k and int(r[2]) == 2}
    assert set(nonempty) <= set(got)
    assert all(g in (1, 2, 3, 4) for g in got.values())
    for b in nonempty:
        assert st["variant"][b] == got[b]
