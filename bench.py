"""LETKF analysis benchmark on one MI355X.

Workload: the k=64 clustered-obs cycle -- grid 360x180x32 (ps + u, v, T:
n = 97 per column), k = 64, 500,000 obs rows (70% on 40 great-circle swaths,
20% at 1,000 station columns, 10% uniform), GC localisation c = 300 km
(support 600 km), one local analysis per column: N_A = 64,800.

A step is one analysis cycle with every input already in HBM: the obs-site
index build (set_obs_device) plus the analysis (selection, obs-space
products, FP64 eigensolve + transform, ensemble update).  The mixed-precision
form is the headline: BF16 split products with FP32 accumulation, FP64
eigensolve and inverse square root, and the perturbations stored in FP32
with the update on f32 MFMA (analyze_device_f32).  The plain FP64 form
(analyze_device, FP64 throughout) is timed beside it.

Prints ONE JSON line (rank 0).  --gpus N > 1 is refused: the engine runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dida-thhpca_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector and FP64 matrix peak
PEAK_BF16_TFLOPS = 2500.0   # dense bf16 MFMA
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


KERNEL_NAMES = {"select": "k_select", "products": "k_products", "tridiag": "k_tridiag",
                "ql": "k_ql_log", "apply": "k_vec", "update": "k_update"}


def kernel_work(name, stats, cfg, n_an, mixed):
    """Algorithmic work of one cycle's launch of a kernel, as the cycle's
    flop and byte counts F and B are defined for this workload:
    (flops, bytes, peak, bound, unit).

    products: the symmetric SYRK k(k+1) p plus Yb^T R^-1 d 2 k p flops per
              analysis (p local obs rows), on MFMA: the BF16 peak in the split
              mode (the three-term split's extra products are not counted),
              FP64 otherwise;
    apply:    T = V f V^T (2 k^3) and overline_w = Pa b (4 k^2) per analysis,
              FP64 MFMA; the eigendecomposition itself (tridiag, QL, the
              eigenvector accumulation) is excluded from F and reported as a
              time share;
    update:   Xa = Xb trans, 2 k^2 m flops, HBM-bound: Xb read and Xa written
              (2 s_x k m bytes, s_x = 4 for the FP32 state of the mixed form,
              8 for FP64) plus the means (16 m bytes) per analysis;
    select, tridiag, ql: no F (index work, eigendecomposition).
    The FP64 vector and FP64 matrix peaks of MI355X are equal (78.6 TFLOP/s)."""
    k, m = cfg.k, cfg.n
    if name == "products":
        p = stats["total_rows"]
        fl = k * (k + 1) * p + 2 * k * p
        return fl, 0, (PEAK_BF16_TFLOPS if mixed == 2 else PEAK_FP64_TFLOPS), "mfma", "TFLOP/s"
    if name == "apply":
        return n_an * (2 * k ** 3 + 4 * k ** 2), 0, PEAK_FP64_TFLOPS, "mfma", "TFLOP/s"
    if name == "update":
        sx = 4 if mixed else 8
        return (n_an * 2 * k ** 2 * m, n_an * (2 * sx * k * m + 16 * m), PEAK_HBM_GBS / 1000.0,
                "hbm", "TB/s")
    return 0, 0, PEAK_FP64_TFLOPS, "none", "TFLOP/s"


def cycle_work(stats, cfg, n_an, n_obs_rows, mixed):
    """F and B of one cycle: F = sum_i [k(k+1) p_i + 2k p_i + 2k^3 + 4k^2 + 2k^2 m_i]
    (the eigendecomposition excluded), B = 2 s_x k N_state + 16 N_state +
    N_obs (s_y k + 40)."""
    k, m = cfg.k, cfg.n
    p = stats["total_rows"]
    F = k * (k + 1) * p + 2 * k * p + n_an * (2 * k ** 3 + 4 * k ** 2 + 2 * k ** 2 * m)
    n_state = cfg.nlat * cfg.nlon * m
    sx = 4 if mixed else 8
    sy = 4 if mixed else 8
    B = 2 * sx * k * n_state + 16 * n_state + n_obs_rows * (sy * k + 40)
    return F, B


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "letkf_c2_k64_traffic.json")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a device-to-device copy of known size), when that
    summary was taken on this workload; else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes")


def roofline(km, stats, cfg, n_an, n_obs_rows, mixed, ms_step, workload=None):
    """The dominant kernel (largest measured time) against its roofline, and
    the whole cycle's F and B against the peaks, with the eigensolver's time
    share (its work is not in F)."""
    cands = [x for x in ("select", "products", "tridiag", "ql", "apply", "update")
             if km.get(x, 0.0) > 0]
    dom = max(cands, key=lambda x: km[x])
    fl, by, peak, bound, unit = kernel_work(dom, stats, cfg, n_an, mixed)
    t = km[dom] * 1e-3
    amount = by if bound == "hbm" else fl
    achieved = amount / t / 1e12
    out = dict(bound=bound, kernel=KERNEL_NAMES[dom], achieved=round(achieved, 4), peak=peak,
               unit=unit, frac=round(achieved / peak, 5),
               traffic=measured_traffic(KERNEL_NAMES[dom], workload),
               ms_per_launch=round(km[dom], 4),
               kernel_ms={KERNEL_NAMES[x]: round(km[x], 4) for x in cands})
    out["bytes_per_launch" if bound == "hbm" else "flops_per_launch"] = amount
    per = {}
    for x in cands:
        f, b, pk, bd, u = kernel_work(x, stats, cfg, n_an, mixed)
        if bd == "none" or km[x] <= 0:
            continue
        a = (b if bd == "hbm" else f) / (km[x] * 1e-3) / 1e12
        per[KERNEL_NAMES[x]] = dict(achieved=round(a, 4), unit=u, frac=round(a / pk, 5))
    out["per_kernel"] = per
    F, B = cycle_work(stats, cfg, n_an, n_obs_rows, mixed)
    tc = ms_step * 1e-3
    eig_ms = km.get("tridiag", 0.0) + km.get("ql", 0.0) + km.get("apply", 0.0)
    out["cycle"] = dict(
        flops=F, bytes=B, ms=round(ms_step, 3),
        tflops=round(F / tc / 1e12, 3), tbps=round(B / tc / 1e12, 4),
        frac_fp64_peak=round(F / tc / 1e12 / PEAK_FP64_TFLOPS, 4),
        frac_hbm=round(B / tc / 1e12 / (PEAK_HBM_GBS / 1000.0), 4),
        eigensolver_ms=round(eig_ms, 3),
        eigensolver_share=round(eig_ms / max(sum(km[x] for x in cands), 1e-9), 4))
    return out


def serving_leg(torch, dev, args, precision):
    """p50/p99 job latency of the serving loop at fractions of its saturation rate."""
    from dida_letkf import serving
    pools = serving.Pools(torch, dev, args.seed, serving.GRIDS, serving.OBS_RANGE[1])
    n_sat = args.serve_sat_jobs or args.serve_jobs
    sat_jobs = serving.draw_jobs(n_sat, None, args.seed + 1)
    # build every pool before timing (backgrounds and obs pools are inputs)
    for j in sat_jobs:
        pools.background(j.grid, j.k)
        pools.obs_pool(j.k, serving.GRIDS[j.grid][2])
    torch.cuda.synchronize()
    log(f"serving: pools ready, {torch.cuda.memory_allocated() / 2**30:.1f} GiB allocated")
    # untimed: one small job per ensemble size first, so that the timed runs
    # do not include the first launch of each k's kernels
    warm = [serving.JobSpec(i=i, arrival=0.0, grid=0, k=k, n_obs=int(serving.OBS_RANGE[0]))
            for i, k in enumerate(serving.KS)]
    for j in warm:
        pools.background(j.grid, j.k)
        pools.obs_pool(j.k, serving.GRIDS[j.grid][2])
    serving.run(torch, warm, pools, precision=precision, seed=args.seed + 7, log=log)
    sat = serving.run(torch, sat_jobs, pools, precision=precision, seed=args.seed, log=log)
    sat_rate = sat.n_jobs / sat.makespan
    out = dict(saturation=sat.summary(), runs=[])
    log(f"serving saturation: {out['saturation']}")
    for frac in [float(x) for x in args.serve_rates.split(",") if x]:
        jobs = serving.draw_jobs(args.serve_jobs, frac * sat_rate, args.seed + 1)
        for j in jobs:
            pools.background(j.grid, j.k)
            pools.obs_pool(j.k, serving.GRIDS[j.grid][2])
        torch.cuda.synchronize()
        r = serving.run(torch, jobs, pools, precision=precision, seed=args.seed + 3, log=log)
        r.rate = frac * sat_rate
        summ = r.summary()
        summ["load"] = frac
        out["runs"].append(summ)
        log(f"serving at {frac:.0%} of saturation: {summ}")
    out["workload"] = (f"Poisson stream, grid from {serving.GRIDS} (nlat, nlon, nz), k from "
                       f"{serving.KS}, obs rows log-uniform in {serving.OBS_RANGE}, "
                       f"{serving.CYCLES} cycles per job, c={serving.LOCAL_DIST_KM:g} km, "
                       f"precision {precision}")
    return out


BUCKET_NAMES = ["0", "1-256", "257-512", "513-1024", "1025-2048", "2049-4096", "4097-8192",
                ">8192"]
VARIANT_NAMES = {1: "dispatch order, 1 row group/step", 2: "dispatch order, 2 row groups/step",
                 3: "persistent queue, 1 row group/step", 4: "persistent queue, 2 row groups/step"}


def products_buckets(st, nob, products_ms):
    """The products schedule's length buckets of p, from the device: analyses,
    share of the rows (the products' work), the variant the tuner chose, and,
    from a step with in-kernel timing, each bucket's share of the workers'
    time and the workers' busy fraction over their launch spans."""
    rows = max(sum(st["bucket_rows"]), 1)
    wave = sum(st["bucket_wave_ms"])
    out = []
    for b, nm in enumerate(BUCKET_NAMES):
        if not st["bucket_analyses"][b]:
            continue
        out.append({"p": nm, "analyses": st["bucket_analyses"][b],
                    "rows_share": round(st["bucket_rows"][b] / rows, 4),
                    "variant": VARIANT_NAMES.get(st["variant"][b], st["variant"][b]),
                    "worker_ms": round(st["bucket_wave_ms"][b], 3),
                    "time_share": round(st["bucket_wave_ms"][b] / wave, 4) if wave else None})
    nob = np.asarray(nob, np.int64) if nob is not None else None
    return {"schedule": ("length buckets, longest first, one launch per run of buckets "
                         "sharing a variant" if st["schedule"] else "column order"),
            "p_mean": round(float(nob.mean()), 1) if nob is not None else None,
            "p_max": int(nob.max()) if nob is not None else None,
            "buckets": out,
            "load_balance": round(st["products_balance"], 4),
            "load_balance_note": ("summed busy time of the products' wavefront groups over "
                                  "the groups the device holds at once x launch span (1: no "
                                  "slot idle from the first start to the last end)"),
            "products_ms": round(products_ms, 3)}


def make_line(args, cfg, n_a, n_obs_rows, res, acc, cpu):
    """The one JSON line: the headline (mixed) mode, the FP64 form beside it."""
    from dida_letkf.config import PREC_FP64
    head = res[args.mixed]
    st = head["stats"]
    workload = (f"{args.config}: {cfg.nlon}x{cfg.nlat}x{cfg.nz} grid, n={cfg.n}, k={cfg.k}, "
                f"{n_obs_rows} clustered obs rows, c={cfg.local_dist:g} km, N_A={n_a} analyses/cycle")
    roof = roofline(head["kernel_ms"], st, cfg, n_a, n_obs_rows, args.mixed, head["ms_step"],
                    workload)
    line = {
        "metric": "LETKF local analyses/sec at k=64 (mixed precision; FP64 form beside it)",
        "value": round(head["value"], 2),
        "unit": "local analyses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("bf16x3" if args.mixed == 2 else "f32") + "+f64" +
                 ("+f32state" if args.mixed_state == "f32" else ""),
        "data": "synthetic",
        "config": {"workload": workload,
                   "n_analyses": n_a, "k": cfg.k, "obs_rows": n_obs_rows,
                   "local_obs_rows_total": st["total_rows"], "local_obs_rows_max": st["max_rows"],
                   "rows_over_10000": st["rows_over_10000"],
                   "precision": (("BF16x3" if args.mixed == 2 else "FP32") + " MFMA obs-space "
                                 "products with FP32 accumulation, FP64 eigensolve and inverse "
                                 "square root, " + ("FP32 perturbation state, update on f32 MFMA"
                                                    if args.mixed_state == "f32" else
                                                    "FP64 state and update")),
                   "parallelism": "1 GPU",
                   "products_schedule": args.schedule,
                   "timed_steps": ("hipGraph replay of the call" if args.graph
                                   else "eager launches")},
        "kernel_ms": head["kernel_ms"],
        "kernel_ms_timed_steps": head.get("kernel_ms_overlapped"),
        "kernel_ms_note": ("HIP-event time of each kernel over the whole batch, from two extra "
                           "steps with the eigensolver on one stream"
                           if "kernel_ms_overlapped" in head else
                           "HIP-event time of each kernel over the whole batch"),
        "roofline": roof,
        "products_buckets": (products_buckets(head["bucket_stats"], head.get("nobsl"),
                                              head["kernel_ms"]["products"])
                             if head.get("bucket_stats") else None),
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    if PREC_FP64 in res:
        r64 = res[PREC_FP64]
        line["fp64_value"] = round(r64["value"], 2)
        line["fp64_ms_per_step"] = round(r64["ms_step"], 3)
        line["fp64_kernel_ms"] = r64["kernel_ms"]
        line["mixed_over_fp64"] = round(head["value"] / r64["value"], 3)
    return line


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--k", type=int, default=0, help="ensemble size (0: the config's)")
    ap.add_argument("--mixed", type=int, default=2, help="precision of the headline mode (1 fp32, 2 bf16x3)")
    ap.add_argument("--mixed-state", choices=("f32", "f64"), default="f32",
                    help="storage of the perturbations in the headline mode (the update runs "
                         "on f32 MFMA for f32)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=8,
                    help="every Nth column of each rank's block for the CPU baseline")
    ap.add_argument("--parity-cols", type=int, default=48)
    ap.add_argument("--serve-jobs", type=int, default=24,
                    help="jobs per arrival rate in the serving leg (0: no serving leg)")
    ap.add_argument("--serve-rates", default="0.8",
                    help="arrival rates as fractions of the measured saturation rate")
    ap.add_argument("--serve-sat-jobs", type=int, default=12,
                    help="jobs of the closed burst that measures saturation (0: = serve-jobs)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the accuracy checks and the local-obs histogram")
    ap.add_argument("--schedule", choices=("queue", "column"), default="queue",
                    help="products over length buckets on persistent work queues, or one "
                         "launch in column order")
    ap.add_argument("--tune", type=int, default=1,
                    help="1: tune the products variant per length bucket before the timed steps")
    ap.add_argument("--graph", action="store_true",
                    help="timed steps replay a captured hipGraph of the call")
    ap.add_argument("--eig-one-stream", action="store_true",
                    help="time the steps with the eigensolver on one stream (transform_kernel "
                         "2; the default overlaps two half-batches on two streams)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device three times (the calibration "
                         "kernel of the HBM traffic counters) and exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported "
              f"(local analyses are independent: the path would shard by latitude bands, "
              f"one band and its obs halo per GPU)", file=sys.stderr)
        sys.exit(2)
    import torch
    from dida_letkf import synth, engine
    from dida_letkf.config import PREC_FP64

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    if args.copy_probe:
        n = 1 << 27
        src = torch.full((n,), 1.5, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_probe_bytes=8 * n, launches=3)), flush=True)
        return
    if args.serve_only:
        out = dict(metric="LETKF serving job latency (p50/p99)",
                   serving=serving_leg(torch, dev, args, args.mixed))
        if not args.no_fp64:
            out["serving_fp64"] = serving_leg(torch, dev, args, PREC_FP64)
        print(json.dumps(out), flush=True)
        return
    t0 = time.time()
    inst = synth.generate(args.config, seed=args.seed, k=args.k or None)
    cfg = inst.cfg
    log(f"generated {args.config}: grid {cfg.nlat}x{cfg.nlon}x{cfg.nz} n={cfg.n} k={cfg.k} "
        f"obs rows={inst.obs.n_rows} sites={inst.obs.n_sites} in {time.time() - t0:.1f}s")
    N_A = cfg.n_analyses()
    o = inst.obs
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d = dict(xb=T(inst.xb), axb=T(inst.axb), lat=T(o.site_lat), lon=T(o.site_lon),
             rp=T(o.site_row_ptr.astype(np.int32)), yb=T(o.yb), dep=T(o.dep), rinv=T(o.rinv),
             vco=T(o.vcoord))
    if args.mixed_state == "f32":
        d["xb32"] = d["xb"].to(torch.float32)
    stream = torch.cuda.current_stream()

    def make_engine(prec, transform_kernel=0, keep=False):
        import copy
        c = copy.copy(cfg)
        c.precision = prec
        eng = engine.LetkfEngine(c, transform_kernel=transform_kernel, keep_transforms=keep)
        eng.set_stream(stream.cuda_stream)
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_option(engine.OPT_SCHEDULE, 1 if args.schedule == "queue" else 0)
        return eng

    def step(eng, xa, axa):
        eng.set_obs_device(o.n_sites, d["lat"].data_ptr(), d["lon"].data_ptr(), d["rp"].data_ptr(),
                           o.n_rows, d["yb"].data_ptr(), d["dep"].data_ptr(), d["rinv"].data_ptr(),
                           d["vco"].data_ptr())
        if xa.dtype == torch.float32:
            return eng.analyze_device_f32(d["xb32"].data_ptr(), d["axb"].data_ptr(),
                                          xa.data_ptr(), axa.data_ptr())
        return eng.analyze_device(d["xb"].data_ptr(), d["axb"].data_ptr(), xa.data_ptr(),
                                  axa.data_ptr())

    def run_mode(prec, transform_kernel=0, steps=None, warmup=None, counts=False, graph=False,
                 bucket_timing=False):
        steps = args.steps if steps is None else steps
        warmup = args.warmup if warmup is None else warmup
        eng = make_engine(prec, transform_kernel, keep=counts)
        f32 = prec != PREC_FP64 and args.mixed_state == "f32"
        xa = torch.zeros_like(d["xb32"] if f32 else d["xb"])
        axa = torch.zeros_like(d["axb"])
        if args.tune and args.schedule == "queue":
            # the tuner measures the products variants per length bucket on
            # the first call (untimed) and caches its choice
            eng.set_option(engine.OPT_TUNE, 1)
            step(eng, xa, axa)
        if graph:
            eng.set_option(engine.OPT_GRAPH, 1)
        for _ in range(warmup):
            step(eng, xa, axa)
        torch.cuda.synchronize()
        k_ms = dict(select=0.0, products=0.0, transform=0.0, tridiag=0.0, ql=0.0, apply=0.0,
                    update=0.0)
        ts = time.perf_counter()
        flagged = 0
        for _ in range(steps):
            flagged += step(eng, xa, axa)
            st = eng.stats
            k_ms["select"] += st.ms_select
            k_ms["products"] += st.ms_products
            k_ms["transform"] += st.ms_transform
            k_ms["tridiag"] += st.ms_tridiag
            k_ms["ql"] += st.ms_ql
            k_ms["apply"] += st.ms_apply
            k_ms["update"] += st.ms_update
        torch.cuda.synchronize()
        el = time.perf_counter() - ts
        stats = eng.stats.as_dict()
        nob = eng.local_obs_counts() if counts else None
        bstats = None
        if bucket_timing and args.schedule == "queue":
            # one more (untimed) step with the products timed per bucket in the kernels
            eng.set_option(engine.OPT_GRAPH, 0)
            eng.set_option(engine.OPT_TIMING, 1)
            step(eng, xa, axa)
            bstats = eng.stats.as_dict()
        eng.close()
        for kk in k_ms:
            k_ms[kk] /= max(steps, 1)
        return dict(elapsed=el, ms_step=1e3 * el / steps, value=N_A * steps / el,
                    kernel_ms=k_ms, stats=stats, flagged=flagged, xa=xa, axa=axa, nobsl=nob,
                    bucket_stats=bstats)

    # timed steps: the eigensolver overlaps two half-batches on two streams
    # (transform_kernel 0), so its kernels are timed one at a time in a short
    # extra leg with the eigensolver on one stream (transform_kernel 2)
    tk = 2 if args.eig_one_stream else 0
    res = {}
    modes = [args.mixed] + ([] if args.no_fp64 else [PREC_FP64])
    for prec in modes:
        r = run_mode(prec, tk, graph=args.graph)
        if tk == 0 or args.graph:
            r["kernel_ms_overlapped"] = r["kernel_ms"]
            rk = run_mode(prec, 2, steps=2, warmup=1, bucket_timing=prec == args.mixed)
            r["kernel_ms"] = rk["kernel_ms"]
            r["bucket_stats"] = rk["bucket_stats"]
        if prec == args.mixed and not args.no_accuracy:
            # local obs counts of the cycle (keep_transforms also writes every
            # trans, so this step is not timed)
            r["nobsl"] = run_mode(prec, 2, steps=1, warmup=0, counts=True)["nobsl"]
        res[prec] = r
        log(f"precision {prec}: {r['value']:.4g} analyses/s, {r['ms_step']:.2f} ms/step, "
            f"kernels (one stream) {r['kernel_ms']}, rows {r['stats']['total_rows']} "
            f"max {r['stats']['max_rows']}")

    # ---- accuracy: FP64 vs oracle on sampled columns; mixed vs FP64 increments
    import orc
    acc = {}
    rng = np.random.default_rng(1)
    if args.no_accuracy:
        args.parity_cols = 0
    n_cols = cfg.nlat * cfg.nlon        # horizontal columns (cube_n = 1); N_A counts
    cols = np.sort(rng.choice(n_cols, min(args.parity_cols, n_cols),   # vertical chunks too
                              replace=False)).astype(np.int32)
    if cols.size:
        ref = orc.letkf_compute(inst, columns=cols, nthreads=min(16, os.cpu_count() or 1))
        ref_full = ref["xa"] + ref["axa"][..., None]
        lon_i, lat_i = cols // cfg.nlat, cols % cfg.nlat
    if PREC_FP64 in res and cols.size:
        g64 = (res[PREC_FP64]["xa"] + res[PREC_FP64]["axa"][..., None]).cpu().numpy()
        diff = np.abs(g64[lon_i, lat_i] - ref_full[lon_i, lat_i]).max()
        acc["fp64_vs_oracle_max_rel"] = float(diff / np.abs(ref_full[lon_i, lat_i]).max())
        gm = (res[args.mixed]["xa"].double() + res[args.mixed]["axa"][..., None])
        g64t = res[PREC_FP64]["xa"] + res[PREC_FP64]["axa"][..., None]
        xbf = d["xb"] + d["axb"][..., None]
        per_var = {}
        vs = cfg.var_start() - 1
        names = ["ps", "u", "v", "T"][: cfg.n_var]
        for v in range(cfg.n_var):
            nl = 1 if cfg.var_lev()[v] == 1 else cfg.nz
            sl = slice(int(vs[v]), int(vs[v]) + nl)
            num = torch.sqrt(torch.mean((gm[:, :, sl] - g64t[:, :, sl]) ** 2))
            den = torch.sqrt(torch.mean((g64t[:, :, sl] - xbf[:, :, sl]) ** 2))
            per_var[names[v]] = float(num / den)
        acc["mixed_increment_rms_rel_err"] = per_var
    log(f"accuracy: {acc}")

    # ---- CPU baseline: the reference's step on the host cores (letkf_baseline.c)
    cpu = None
    if not args.no_cpu:
        # 16 threads: the GPU box's CPU share for one GPU
        ncore = min(16, os.cpu_count() or 1)
        b = orc.baseline(inst, nthreads=ncore, sample=args.cpu_sample)
        n_an_cpu = b["n_analyses"] * (N_A // (cfg.nlat * cfg.nlon))
        cpu = dict(value=float(n_an_cpu / b["wall_s"]), unit="local analyses/s", cores=ncore,
                   kind="port",
                   sample=(f"every {args.cpu_sample}th column of each rank's block of the same "
                           f"cycle ({n_an_cpu} analyses, {b['wall_s']:.1f} s wall to the slowest "
                           f"rank); {b['split'][0]}x{b['split'][1]} (lat x lon) rank blocks as "
                           f"MPI_Dims_create({ncore}, 2), each scanning the obs of its "
                           f"comm_info_init neighbour blocks "
                           f"({b['n_candidates'] / max(b['n_analyses'], 1):.0f} sites per "
                           f"analysis); letkf_core and xb_2_xa on {b['provider']}, FP64"))
        log(f"cpu baseline: {cpu}")

    line = make_line(args, cfg, N_A, o.n_rows, res, acc, cpu)
    if args.serve_jobs > 0:
        del d
        torch.cuda.empty_cache()
        line["serving"] = serving_leg(torch, dev, args, args.mixed)
        if not args.no_fp64:
            line["serving_fp64"] = serving_leg(torch, dev, args, PREC_FP64)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
