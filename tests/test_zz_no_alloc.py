"""dida_letkf_core_batched_noalloc (no device allocation, fixed scratch)
against dida_letkf_core_batched on the same inputs: results bit for bit equal
(the same kernels on the same data, one stream), on two calls in a row with
the second smaller, in every precision mode and kernel width, and with the
eigensolver's list pass forced."""
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


def _cases(k, sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for p in sizes:
        Y = rng.standard_normal((k, p))
        Y -= Y.mean(axis=0, keepdims=True)
        out.append((Y, (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32),
                    rng.standard_normal(p), rng.uniform(0.0, 1.0, p)))
    return [[c[i] for c in out] for i in range(4)]


def _both(args, **kw):
    from dida_letkf import engine
    a = engine.letkf_core_batched(*args, **kw)
    b = engine.letkf_core_batched(*args, noalloc=True, **kw)
    return a, b


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("k", [20, 64, 128])
def test_noalloc_matches_allocating_entry(k, prec):
    big = _cases(k, [0, 1, 5, k - 1, k, 3 * k, 700, 33] * 3, seed=k + prec)
    small = _cases(k, [2, 300, 0, k + 1], seed=100 + k + prec)
    for args in (big, small):           # two calls in a row, the second smaller
        (t0, w0, s0), (t1, w1, s1) = _both(args, inflation_factor=1.07, precision=prec)
        assert np.array_equal(s0, s1) and (s0 == 0).all()
        assert np.array_equal(t0, t1)
        assert np.array_equal(w0, w1)


def test_noalloc_list_pass():
    """The eigensolver's list pass (forced from the host) inside the fixed
    scratch gives the same results as in the allocating entry point."""
    args = _cases(64, [10, 64, 200, 500, 1000], seed=7)
    old = os.environ.get("DIDA_ROT_CAP")
    os.environ["DIDA_ROT_CAP"] = "16"
    try:
        (t0, w0, s0), (t1, w1, s1) = _both(args, precision=0)
    finally:
        if old is None:
            os.environ.pop("DIDA_ROT_CAP", None)
        else:
            os.environ["DIDA_ROT_CAP"] = old
    assert (s0 == 0).all() and np.array_equal(s0, s1)
    assert np.array_equal(t0, t1) and np.array_equal(w0, w1)


def test_noalloc_rejects_oversize():
    from dida_letkf import engine
    args = _cases(128, [10] * 200, seed=9)          # ~2 MB of scratch per analysis
    with pytest.raises(engine.DidaError, match="scratch"):
        engine.letkf_core_batched(*args, noalloc=True)
    # and the entry point still works afterwards
    small = _cases(128, [10, 200], seed=10)
    (t0, w0, _), (t1, w1, _) = _both(small)
    assert np.array_equal(t0, t1) and np.array_equal(w0, w1)
