"""GPU tests of the eigensolver's failure paths and of the reduced-precision
guard, through the C ABI.

* Flag paths: the first QL pass flags a matrix whose rotation log fills up or
  whose QL does not converge in its iteration limit; a list pass redoes it
  with a longer log and more iterations.  The limits are lowered from the
  host (DIDA_ROT_CAP, DIDA_QL_MAXITER, DIDA_RERUN_ROT_CAP,
  DIDA_RERUN_MAXITER), so every kernel still runs to its normal end.
* Guard (letkf_core_mpal.F90:113-118, Sunway letkf_core_half_master.c:157):
  at k <= 64 the BF16x3 products drop the lo*lo term, so the reduced A is
  A - sum_j lo_j lo_j^T, which has negative eigenvalues when A is rank
  deficient.  With w = 1 and integer Yb of at most 12 significant bits every
  product and partial sum is exact in FP32, so the reduced A is known exactly
  on the CPU and the guarded transform of it is checked at 1e-10.
* Half-batch pipeline (two streams, batches of 2048 analyses or more) against
  the one-stream eigensolver bit for bit and against the oracle.
"""
import os

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


class _Env:
    """Sets environment variables for the duration of a with-block."""

    def __init__(self, **kw):
        self.kw = {k: str(v) for k, v in kw.items()}
        self.old = {}

    def __enter__(self):
        for k, v in self.kw.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = v

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _dense_instance(k, nlat, nlon, seed):
    """Every column has local obs (dense uniform obs, GC support ~ 1800 km)."""
    from dida_letkf import synth
    return synth.generate("c0", seed=seed, k=k, nlat=nlat, nlon=nlon, n_obs=1500,
                          local_dist=900.0)


def _run(inst, **kw):
    from dida_letkf import engine
    return engine.letkf_compute_nooverlap(inst, keep_transforms=True, **kw)


# ---------------------------------------------------------------------------
# flag paths
# ---------------------------------------------------------------------------

def _column(cfg, a):
    """(lon, lat) index in the haloed [lon][lat] state of horizontal analysis a."""
    return a // cfg.nlat + cfg.halo, a % cfg.nlat + cfg.halo


@pytest.mark.parametrize("k", [20, 64, 128])
@pytest.mark.parametrize("hook", ["DIDA_ROT_CAP", "DIDA_QL_MAXITER"])
def test_list_pass_redoes_flagged(k, hook):
    """A first pass that flags (nearly) every matrix (rotation log of 16
    entries, or a single QL iteration); the list pass's results equal the
    oracle's."""
    import orc
    inst = _dense_instance(k, 8, 16, seed=40 + k)          # 128 analyses <= list capacity
    ref = orc.letkf_compute(inst, want_trans=True)
    val = 16 if hook == "DIDA_ROT_CAP" else 1
    with _Env(**{hook: val}):
        gpu = _run(inst)
    st = gpu["stats"]
    print(k, hook, "rerun", st["n_rerun"], "flagged", st["n_flagged"])
    assert st["n_rerun"] >= 96
    assert gpu["n_flagged"] == 0 and st["n_flagged"] == 0
    assert np.array_equal(gpu["nobsl"], ref["nobsl"])
    assert rel(gpu["xa"], ref["xa"]) < TOL64
    assert rel(gpu["axa"], ref["axa"]) < TOL64
    assert rel(gpu["trans"], ref["trans"]) < TOL64


def _check_split(inst, gpu, ref, n_expected_failed=None):
    """Analyses with NaN transforms kept the background exactly; the others
    match the oracle."""
    cfg = inst.cfg
    failed = np.isnan(gpu["trans"]).all(axis=(1, 2))
    assert failed.sum() == gpu["n_flagged"] == gpu["stats"]["n_flagged"]
    if n_expected_failed is not None:
        assert failed.sum() == n_expected_failed
    assert np.isnan(gpu["overline_w"][failed]).all()
    assert np.isfinite(gpu["trans"][~failed]).all()
    for a in np.flatnonzero(failed):
        jl, il = _column(cfg, a)
        assert np.array_equal(gpu["xa"][jl, il], inst.xb[jl, il])
        assert np.array_equal(gpu["axa"][jl, il], inst.axb[jl, il])
    ok = ~failed
    if ok.any():
        assert rel(gpu["trans"][ok], ref["trans"][ok]) < TOL64
        for a in np.flatnonzero(ok)[:64]:
            jl, il = _column(cfg, a)
            assert rel(gpu["xa"][jl, il], ref["xa"][jl, il]) < 1e-9
    return failed


@pytest.mark.parametrize("k", [20, 128])
def test_failed_twice_keeps_background(k):
    """Both passes flag the matrices with long chases: their state keeps the
    background (xa = xb, axa = axb exactly), trans and overline_w are NaN,
    and the return code counts them."""
    import orc
    inst = _dense_instance(k, 6, 12, seed=50 + k)
    ref = orc.letkf_compute(inst, want_trans=True)
    with _Env(DIDA_ROT_CAP=16, DIDA_RERUN_ROT_CAP=16):
        gpu = _run(inst)
    failed = _check_split(inst, gpu, ref)
    assert failed.sum() >= 0.75 * gpu["stats"]["n_analyses"]


def test_more_flagged_than_list_capacity():
    """288 analyses, (nearly) all flagged, at most 256 redone by the list pass
    (k <= 64): the rest keep the background, every redone one matches the
    oracle."""
    import orc
    inst = _dense_instance(32, 12, 24, seed=61)
    ref = orc.letkf_compute(inst, want_trans=True)
    with _Env(DIDA_ROT_CAP=16):
        gpu = _run(inst)
    n = gpu["stats"]["n_analyses"]
    n_rerun = gpu["stats"]["n_rerun"]
    assert n == 288 and n_rerun > 256
    _check_split(inst, gpu, ref, n_expected_failed=n_rerun - 256)


def test_batched_core_list_pass():
    """dida_letkf_core_batched (k = 100, mixed sizes) with the first pass's
    log forced full."""
    import orc
    from dida_letkf import engine
    rng = np.random.default_rng(5)
    k = 100
    cases = []
    for p in (0, 1, 50, 99, 100, 400):
        Y = rng.standard_normal((k, p))
        cases.append((Y, (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32),
                      rng.standard_normal(p), rng.uniform(0, 1, p)))
    with _Env(DIDA_ROT_CAP=16):
        t, w, st = engine.letkf_core_batched(*[[c[i] for c in cases] for i in range(4)],
                                             inflation_factor=1.1)
    assert (st == 0).all()
    for i, c in enumerate(cases):
        t_o, w_o, _ = orc.letkf_core(*c, inflation_factor=1.1)
        assert rel(t[i], t_o) < TOL64, i
        assert np.abs(w[i] - w_o).max() < TOL64 * max(1.0, np.abs(w_o).max()), i


# ---------------------------------------------------------------------------
# reduced-precision guard
# ---------------------------------------------------------------------------

def _bf16_rne(x):
    """float32 -> bfloat16 (round to nearest even), as float64 values."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _guard_case(k, p, seed):
    rng = np.random.default_rng(seed)
    # integers of at most 12 significant bits, away from zero
    Y = rng.integers(1024, 2048, size=(k, p)) * rng.choice([-1, 1], size=(k, p))
    Y = Y.astype(np.float64)
    dep = rng.integers(-3, 4, size=p).astype(np.float64)
    rinv = np.ones(p, np.float32)
    rloc = np.ones(p)
    return Y, rinv, dep, rloc


@pytest.mark.parametrize("k", [8, 40])
def test_bf16x3_guard_fires(k):
    """k <= 64: the two-term split (hi.hi + hi.lo + lo.hi) of 12-bit integers."""
    import orc
    from dida_letkf import engine
    p = 3
    Y, rinv, dep, rloc = _guard_case(k, p, seed=k)
    hi = _bf16_rne(Y)
    lo = _bf16_rne(Y - hi)
    assert np.array_equal(hi + lo, Y)                  # the split is exact
    A_red = hi @ hi.T + hi @ lo.T + lo @ hi.T          # exact: integers below 2^24
    b_red = Y @ dep
    lam = np.linalg.eigvalsh(A_red)
    lam_min = lam.min()
    assert lam_min < -10.0                             # far below -(k-1)/rho's scale
    t_g, w_g, st = engine.letkf_core_batched([Y], [rinv], [dep], [rloc], inflation_factor=1.0,
                                             precision=2)
    assert st[0] == 0
    t_o, w_o, _ = orc.letkf_transform(A_red, b_red, guard=True)
    t_u, _, _ = orc.letkf_transform(A_red + 0.0, b_red, guard=False)
    assert np.isfinite(t_o).all()
    # unguarded, lam_min + (k-1)/rho < 0 and the inverse square root is NaN
    assert not np.isfinite(t_u).all()
    # Two FP64 eigensolvers agree on the shifted small eigenvalues (~(k-1)/rho
    # after the shift) to about eps * |A|: the bar scales with |A| / (k-1),
    # which the 2^-16 ratio of lo*lo to y*y makes ~1e5..1e7 here
    tol = max(TOL64, np.finfo(float).eps * lam.max() / (k - 1))
    print(k, "lam", lam_min, lam.max(), "tol", tol, "err", rel(t_g[0], t_o))
    assert tol < 2e-9
    assert rel(t_g[0], t_o) < tol
    assert np.abs(w_g[0] - w_o).max() < tol * max(1.0, np.abs(w_o).max())


def test_bf16x3_three_terms_exact_k100():
    """k = 65..128 splits each value in three BF16 terms (24 bits): 12-bit
    integers are then represented exactly, A is the exact positive
    semi-definite product, and the guarded result is the FP64 oracle's."""
    import orc
    from dida_letkf import engine
    Y, rinv, dep, rloc = _guard_case(100, 3, seed=100)
    t_g, w_g, st = engine.letkf_core_batched([Y], [rinv], [dep], [rloc], precision=2)
    t_o, w_o, _ = orc.letkf_core(Y, rinv, dep, rloc, guard=True)
    lam_max = np.linalg.eigvalsh(Y @ Y.T).max()
    tol = max(TOL64, np.finfo(float).eps * lam_max / 99)
    assert st[0] == 0
    assert rel(t_g[0], t_o) < tol
    assert np.abs(w_g[0] - w_o).max() < tol * max(1.0, np.abs(w_o).max())


@pytest.mark.parametrize("k", [8, 100])
def test_fp32_guard_on_exact_products(k):
    """FP32 products of the same integers are exact, so A is positive
    semi-definite with exact zero eigenvalues; the guard then shifts by at
    most rounding and the result is the FP64 oracle's."""
    import orc
    from dida_letkf import engine
    Y, rinv, dep, rloc = _guard_case(k, 3, seed=3 * k)
    t_g, w_g, st = engine.letkf_core_batched([Y], [rinv], [dep], [rloc], precision=1)
    t_o, w_o, _ = orc.letkf_core(Y, rinv, dep, rloc, guard=True)
    assert st[0] == 0
    assert rel(t_g[0], t_o) < 1e-9


# ---------------------------------------------------------------------------
# half-batch pipeline (two streams) at >= 2048 analyses
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("k", [64, 128])
def test_half_batch_pipeline(k):
    import orc
    from dida_letkf import synth
    inst = synth.generate("c0", seed=70 + k, k=k, nlat=48, nlon=48, n_obs=400,
                          local_dist=600.0)
    assert inst.cfg.n_analyses() >= 2048
    two = _run(inst, transform_kernel=0)
    one = _run(inst, transform_kernel=2)
    assert np.array_equal(two["xa"], one["xa"]) and np.array_equal(two["axa"], one["axa"])
    assert np.array_equal(two["trans"], one["trans"])
    ref = orc.letkf_compute(inst, want_trans=True)
    assert two["n_flagged"] == 0
    assert np.array_equal(two["nobsl"], ref["nobsl"])
    assert rel(two["xa"], ref["xa"]) < TOL64 and rel(two["axa"], ref["axa"]) < TOL64
    assert rel(two["trans"], ref["trans"]) < TOL64
