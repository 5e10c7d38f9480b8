"""bench.py's JSON line is assembled on the CPU from measured figures: the
keys the contract requires, and the roofline of the dominant kernel."""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_line_has_contract_keys():
    sys.path.insert(0, ROOT)
    import bench
    from dida_letkf import synth
    cfg = synth.make_config("c2")
    args = argparse.Namespace(mixed=2, mixed_state="f32", steps=5, warmup=1, config="c2",
                              schedule="queue", graph=False)
    km = dict(select=2.8, products=9.6, transform=41.9, tridiag=8.5, ql=9.0, apply=24.0, update=1.5)
    st = dict(total_rows=72_613_192, max_rows=2656, ms_update=1.5, rows_over_10000=0)
    bst = dict(schedule=1, bucket_analyses=[0, 10, 20, 30, 0, 0, 0, 0],
               bucket_rows=[0, 1000, 6000, 21000, 0, 0, 0, 0],
               bucket_wave_ms=[0.0, 1.0, 3.0, 6.0, 0.0, 0.0, 0.0, 0.0],
               variant=[1, 1, 2, 3, 1, 1, 1, 1], products_balance=0.9)
    res = {2: dict(value=1.18e6, ms_step=54.9, kernel_ms=km, stats=st, bucket_stats=bst,
                   nobsl=[100, 300, 700]),
           0: dict(value=1.0e6, ms_step=64.8, kernel_ms=dict(km, products=19.9), stats=st)}
    line = bench.make_line(args, cfg, cfg.n_analyses(), 500_000, res, {}, None)
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in line
    assert line["config"]["rows_over_10000"] == 0
    r = line["roofline"]
    assert r["kernel"] == "k_vec" and r["bound"] == "mfma" and r["unit"] == "TFLOP/s"
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-4
    k, m, n_a = cfg.k, cfg.n, cfg.n_analyses()
    # the eigenvector kernel is priced at T = V f V^T and overline_w only
    fl = n_a * (2 * k**3 + 4 * k**2)
    assert abs(r["achieved"] - fl / 24.0e-3 / 1e12) < 1e-3
    # the cycle's F and B, and the eigensolver's share of the kernel time
    p = st["total_rows"]
    F = k * (k + 1) * p + 2 * k * p + n_a * (2 * k**3 + 4 * k**2 + 2 * k**2 * m)
    n_state = cfg.nlat * cfg.nlon * m
    B = 2 * 4 * k * n_state + 16 * n_state + 500_000 * (4 * k + 40)
    assert r["cycle"]["flops"] == F and r["cycle"]["bytes"] == B
    assert abs(r["cycle"]["eigensolver_share"] - (8.5 + 9.0 + 24.0) / 55.4) < 1e-3
    # an HBM-bound update, when it dominates, is priced in bytes (FP32 state)
    r2 = bench.roofline(dict(update=3.0, apply=1.0), st, cfg, n_a, 500_000, 2, 4.0)
    assert r2["bound"] == "hbm" and r2["unit"] == "TB/s"
    assert abs(r2["achieved"] - n_a * (8 * k * m + 16 * m) / 3.0e-3 / 1e12) < 1e-3
    assert line["mixed_over_fp64"] == round(1.18e6 / 1.0e6, 3)
    pb = line["products_buckets"]
    assert [b["p"] for b in pb["buckets"]] == ["1-256", "257-512", "513-1024"]
    assert [b["rows_share"] for b in pb["buckets"]] == [round(1 / 28, 4), round(6 / 28, 4),
                                                          0.75]
    assert [b["time_share"] for b in pb["buckets"]] == [0.1, 0.3, 0.6]
    assert pb["load_balance"] == 0.9 and pb["p_max"] == 700
    json.dumps(line)


def test_traffic_from_committed_profile():
    """roofline.traffic comes from the committed rocprofv3 --pmc summary, and
    only for the workload it was measured on."""
    sys.path.insert(0, ROOT)
    import bench
    with open(bench.TRAFFIC_FILE) as f:
        t = json.load(f)
    k = t["kernels"]["k_vec"]
    est = (k["fetch_kib"] * t["fetch_factor"] + k["write_kib"] * t["write_factor"]) * 1024
    assert abs(k["bytes"] - est) < 1e-3 * est
    assert bench.measured_traffic("k_vec", t["workload"]) == k["bytes"]
    assert bench.measured_traffic("k_vec", "another workload") is None
    assert bench.measured_traffic("no_such_kernel", t["workload"]) is None
