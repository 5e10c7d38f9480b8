"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py
from the pinned oracle).  CPU: the oracle still reproduces them.  GPU: the
HIP engine reproduces them through the C ABI."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _instance(z):
    from dida_letkf import synth
    from dida_letkf.synth import Instance, Obs
    cfg = synth.make_config("c0", nlat=int(z["grid_lat"].size), nlon=int(z["grid_lon"].size))
    cfg.local_dist = float(z["local_dist"])
    obs = Obs(site_lat=z["site_lat"], site_lon=z["site_lon"], site_row_ptr=z["site_row_ptr"],
              yb=z["yb"], dep=z["dep"], rinv=z["rinv"], vcoord=z["vcoord"],
              row_var=np.zeros(z["dep"].size, np.int32))
    return Instance(cfg=cfg, grid_lat=z["grid_lat"], grid_lon=z["grid_lon"], xb=z["xb"],
                    axb=z["axb"], truth=None, obs=obs)


def test_oracle_reproduces_gc_and_distance_tables(orc):
    z = _load("schr2_latlon.npz")
    gc = np.array([orc.schr2(r) for r in z["ri"]])
    assert np.array_equal(gc, z["schr2"])
    d = np.array([orc.latlon_dist(*q) for q in z["pts"]])
    assert np.abs(d - z["dist"]).max() <= 1e-9 * np.abs(z["dist"]).max()


def test_oracle_reproduces_letkf_core_cases(orc):
    z = _load("letkf_core.npz")
    for i, (k, p) in enumerate(z["cases"]):
        t, w, info = orc.letkf_core(z[f"c{i}_hxens"], z[f"c{i}_rdiaginv"], z[f"c{i}_dep"],
                                    z[f"c{i}_rloc"], float(z["inflation_factor"]))
        assert info == 0
        assert np.abs(t - z[f"c{i}_trans"]).max() < 1e-13, (k, p)
        assert np.abs(w - z[f"c{i}_w"]).max() < 1e-13, (k, p)


def test_oracle_reproduces_c0_decimated(orc):
    z = _load("c0_decimated.npz")
    inst = _instance(z)
    ref = orc.letkf_compute(inst, want_trans=True, nthreads=2)
    assert np.array_equal(ref["nobsl"], z["nobsl"])
    assert np.abs(ref["xa"] - z["xa"]).max() < 1e-12
    assert np.abs(ref["axa"] - z["axa"]).max() < 1e-12
    assert np.abs(ref["overline_w"] - z["overline_w"]).max() < 1e-12


@pytest.mark.gpu
def test_gpu_matches_letkf_core_golden():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dida_letkf import engine
    z = _load("letkf_core.npz")
    for i, (k, p) in enumerate(z["cases"]):
        t, w = engine.letkf_core(z[f"c{i}_hxens"], z[f"c{i}_rdiaginv"], z[f"c{i}_dep"],
                                 z[f"c{i}_rloc"], float(z["inflation_factor"]))
        scale = max(1.0, np.abs(z[f"c{i}_trans"]).max())
        assert np.abs(t - z[f"c{i}_trans"]).max() < 1e-10 * scale, (k, p)
        assert np.abs(w - z[f"c{i}_w"]).max() < 1e-10 * max(1.0, np.abs(z[f"c{i}_w"]).max()), (k, p)


@pytest.mark.gpu
def test_gpu_matches_c0_decimated_golden():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dida_letkf import engine
    z = _load("c0_decimated.npz")
    inst = _instance(z)
    out = engine.letkf_compute_nooverlap(inst, keep_transforms=True)
    assert np.array_equal(out["nobsl"], z["nobsl"])
    for key in ("xa", "axa", "overline_w"):
        ref = z[key]
        assert np.abs(out[key] - ref).max() < 1e-10 * max(1.0, np.abs(ref).max()), key
