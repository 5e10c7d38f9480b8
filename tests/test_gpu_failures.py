"""Failure handling forced from the host, through the C ABI:

* a NaN or Inf in one obs row's Yb, d or R^-1: every analysis whose local
  list holds the row keeps its background (its status is nonzero and it is
  counted in n_flagged); every other analysis gives the same bits as without
  the bad row;
* a device allocation refused at each allocation point of a handle's life
  (DIDA_FAIL_ALLOC_AFTER = 0, 1, 2, ...): the call fails with DIDA_ERR_NOMEM
  and a later call with allocation allowed gives the clean result.

Forced eigensolver non-convergence (QL iteration and log limits) is covered
by test_gpu_flags.py; admission refusal and cancellation by test_gpu_serve.py.
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


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


def _analyze(inst, obs=None, precision=None, transform_kernel=0):
    import copy
    import torch
    from dida_letkf import engine
    cfg = inst.cfg
    if precision is not None:
        cfg = copy.copy(cfg)
        cfg.precision = precision
    with engine.LetkfEngine(cfg, transform_kernel=transform_kernel) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.set_obs(obs if obs is not None else inst.obs)
        xb, axb = _dev(inst.xb), _dev(inst.axb)
        xa, axa = torch.zeros_like(xb), torch.zeros_like(axb)
        rc = eng.analyze_device(xb.data_ptr(), axb.data_ptr(), xa.data_ptr(), axa.data_ptr())
        torch.cuda.synchronize()
        return rc, xa.cpu().numpy(), axa.cpu().numpy(), eng.stats.as_dict()


def _instance(k):
    from dida_letkf import synth
    if k == 128:
        return synth.generate("c3", seed=91, k=128, nlat=10, nlon=20, nz=2, n_obs=6000,
                              local_dist=700.0)
    return synth.generate("c2", seed=92, k=k, nlat=14, nlon=28, nz=2, n_obs=6000,
                          local_dist=700.0)


def _poison(obs, field, value):
    import copy
    bad = copy.deepcopy(obs)
    # a row of a site near the middle of the site list
    r = int(bad.site_row_ptr[bad.n_sites // 2])
    arr = getattr(bad, field)
    if arr.ndim == 2:
        arr[r, 3] = value
    else:
        arr[r] = value
    return bad


CASES = [("yb", np.nan, 64, 2, 0), ("yb", np.inf, 64, 0, 0), ("dep", np.nan, 64, 2, 0),
         ("dep", np.inf, 64, 0, 0), ("rinv", np.inf, 64, 2, 0), ("dep", np.nan, 128, 2, 0),
         ("yb", np.nan, 128, 0, 0), ("dep", np.nan, 32, 0, 1)]


@pytest.mark.parametrize("field,value,k,prec,tk", CASES)
def test_non_finite_obs_keep_background(field, value, k, prec, tk):
    inst = _instance(k)
    print(field, value, "k", k, "precision", prec, "transform kernel", tk, flush=True)
    rc0, xa0, axa0, st0 = _analyze(inst, precision=prec, transform_kernel=tk)
    assert rc0 == 0
    rc, xa, axa, st = _analyze(inst, _poison(inst.obs, field, value), precision=prec,
                               transform_kernel=tk)
    h = inst.cfg.halo
    sl = (slice(h, h + inst.cfg.nlon), slice(h, h + inst.cfg.nlat))
    xb, axb = inst.xb[sl], inst.axb[sl]
    xa, axa, xa0, axa0 = xa[sl], axa[sl], xa0[sl], axa0[sl]
    assert np.all(np.isfinite(xa)) and np.all(np.isfinite(axa))
    same = np.all(xa == xa0, axis=(2, 3)) & np.all(axa == axa0, axis=2)
    if tk == 1:
        # the single-kernel transform uses trans = I: the background up to the
        # re-centring of the update
        kept = (np.abs(xa - xb).max(axis=(2, 3)) < 1e-12) & \
               (np.abs(axa - axb).max(axis=2) < 1e-12 * np.abs(axb).max())
    else:
        kept = np.all(xa == xb, axis=(2, 3)) & np.all(axa == axb, axis=2)
    print("analyses keeping the background", int((kept & ~same).sum()), "flagged", st["n_flagged"])
    assert np.all(same | kept)
    assert rc == st["n_flagged"] == int((~same).sum()) > 0


def test_allocation_refused_at_every_point():
    import torch
    from dida_letkf import engine
    inst = _instance(32)
    rc0, xa0, axa0, _ = _analyze(inst)
    n_fail = 0
    for n in range(0, 200):
        os.environ["DIDA_FAIL_ALLOC_AFTER"] = str(n)
        try:
            _analyze(inst)
        except engine.DidaError as e:
            assert "(-3)" in str(e), str(e)             # DIDA_ERR_NOMEM
            n_fail += 1
            continue
        finally:
            os.environ.pop("DIDA_FAIL_ALLOC_AFTER", None)
        break
    else:
        pytest.fail("every allocation count failed")
    print("allocation points", n_fail)
    assert n_fail > 10
    # allocation allowed again: the clean result
    rc, xa, axa, _ = _analyze(inst)
    assert rc == 0 and np.array_equal(xa, xa0) and np.array_equal(axa, axa0)
    torch.cuda.synchronize()
