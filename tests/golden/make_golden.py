"""Writes the committed golden vectors under tests/golden/.

The reference ships no golden vectors for the LETKF path, so these fixtures
are written by the pinned CPU oracle (oracle/letkf_oracle.c), whose own
tests pin it to closed forms, textbook formulas and LAPACK.  They freeze the
oracle's outputs so that the GPU engine (tests marked gpu) and later
versions of the oracle are checked against fixed data.

  schr2_latlon.npz    GC values on a grid of ri; latlon_dist on random pairs
  letkf_core.npz      12 letkf_core cases, k in {2,3,8,20} x p in {0,1,17,300}
                      (one near-degenerate spectrum among them), inputs and
                      (trans, overline_w)
  c0_decimated.npz    the c0 workload's generator (72x36 grid, k=20) on a 8x4-decimated grid
                      (9 x 9 columns, 250 obs): inputs and (xa, axa,
                      overline_w, nobsl)

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "dida-thhpca_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import orc  # noqa: E402
from dida_letkf import synth  # noqa: E402

CORE_CASES = [(k, p) for k in (2, 3, 8, 20) for p in (0, 1, 17, 300)][:12]


def core_inputs(k, p, seed):
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((k, p))
    if p and k == 8 and p == 17:
        # near-degenerate spectrum: two nearly parallel obs directions
        Y[:, 1] = Y[:, 0] * (1 + 1e-9)
    Y -= Y.mean(axis=0, keepdims=True) if p else 0.0
    rinv = (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32)
    dep = rng.standard_normal(p)
    rloc = rng.uniform(0.0, 1.0, p)
    return Y, rinv, dep, rloc


def main():
    # GC and distance tables
    ri = np.linspace(0.0, 2.2, 221)
    gc = np.array([orc.schr2(r) for r in ri])
    rng = np.random.default_rng(42)
    pts = np.column_stack([rng.uniform(-90, 90, 200), rng.uniform(0, 360, 200),
                           rng.uniform(-90, 90, 200), rng.uniform(0, 360, 200)])
    dist = np.array([orc.latlon_dist(*q) for q in pts])
    np.savez_compressed(os.path.join(HERE, "schr2_latlon.npz"), ri=ri, schr2=gc, pts=pts,
                        dist=dist)

    # letkf_core cases
    out = {}
    for i, (k, p) in enumerate(CORE_CASES):
        Y, rinv, dep, rloc = core_inputs(k, p, 1000 + i)
        trans, w, info = orc.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.05)
        assert info == 0
        out[f"c{i}_hxens"] = Y
        out[f"c{i}_rdiaginv"] = rinv
        out[f"c{i}_dep"] = dep
        out[f"c{i}_rloc"] = rloc
        out[f"c{i}_trans"] = np.asarray(trans)
        out[f"c{i}_w"] = w
    out["cases"] = np.array(CORE_CASES)
    out["inflation_factor"] = np.array(1.05)
    np.savez_compressed(os.path.join(HERE, "letkf_core.npz"), **out)

    # the c0 workload decimated by 8 in lon and 4 in lat: 9 x 9 columns
    inst = synth.generate("c0", seed=3, nlat=9, nlon=9, n_obs=250)
    ref = orc.letkf_compute(inst, want_trans=True, nthreads=1)
    o = inst.obs
    np.savez_compressed(
        os.path.join(HERE, "c0_decimated.npz"),
        grid_lat=inst.grid_lat, grid_lon=inst.grid_lon, xb=inst.xb, axb=inst.axb,
        site_lat=o.site_lat, site_lon=o.site_lon, site_row_ptr=o.site_row_ptr, yb=o.yb,
        dep=o.dep, rinv=o.rinv, vcoord=o.vcoord, k=inst.cfg.k, local_dist=inst.cfg.local_dist,
        xa=ref["xa"], axa=ref["axa"], overline_w=ref["overline_w"],
        nobsl=ref["nobsl"])
    for f in ("schr2_latlon.npz", "letkf_core.npz", "c0_decimated.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
