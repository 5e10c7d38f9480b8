"""Serving workload: a Poisson stream of LETKF analysis jobs of mixed grid,
ensemble size and obs count through the continuous-batching server
(``dida_serve_*``).

Each job draws its grid from GRIDS, k from KS and its obs count log-uniform
in OBS_RANGE, and runs CYCLES analysis cycles with its ensemble resident on
the device.  Latency is arrival -> end of the job's last cycle, on the
server's clock: it includes queueing for admission and any wait of the
client (whose output arrays for queued jobs must exist at submit).

Inputs are synthetic and device-generated, because a job stream of this size
cannot be built through the host-side observation operator in the time of a
benchmark: backgrounds are mean-zero Gaussian perturbations, obs sites follow
the clustered mixture (70% on 40 great-circle swaths 50 km wide, 1,000 station
columns with every level of the three 3-D variables, the rest uniform),
Yb rows are centred Gaussians and R = diag(1 or 4) by variable.  The cost of
an analysis depends on these only through the local obs counts (set by the
site positions and the localisation radius) and k, which follow the workload.
Every job of one (grid, k) shape reads the same background arrays; each cycle
of a job uses its own contiguous range of a shuffled site pool.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import synth
from .config import PREC_BF16X3

GRIDS: List[Tuple[int, int, int]] = [(90, 180, 16), (180, 360, 32), (360, 720, 32)]  # nlat, nlon, nz
KS: List[int] = [20, 32, 64, 128]
OBS_RANGE = (1.0e4, 2.0e6)
CYCLES = 3
LOCAL_DIST_KM = 300.0      # GC half-width c; support 2c = 600 km (clustered-config convention)


@dataclass
class JobSpec:
    i: int
    arrival: float
    grid: int
    k: int
    n_obs: int


@dataclass
class ServeResult:
    n_jobs: int
    rate: Optional[float]
    latency: List[float] = field(default_factory=list)
    queue: List[float] = field(default_factory=list)
    analyses: int = 0
    makespan: float = 0.0
    failed: int = 0
    client_stalls: int = 0          # jobs whose submit waited for output memory
    client_stall_s: float = 0.0     # the waits' total

    def summary(self) -> dict:
        lat = np.asarray(self.latency)
        q = np.asarray(self.queue)
        return dict(
            n_jobs=self.n_jobs,
            arrival_rate_jobs_per_s=self.rate,
            p50_latency_s=float(np.percentile(lat, 50)) if lat.size else None,
            p99_latency_s=float(np.percentile(lat, 99)) if lat.size else None,
            mean_queue_s=float(q.mean()) if q.size else None,
            makespan_s=self.makespan,
            jobs_per_s=self.n_jobs / self.makespan if self.makespan > 0 else None,
            analyses_per_s=self.analyses / self.makespan if self.makespan > 0 else None,
            failed=self.failed,
            client_stalls=self.client_stalls,
            client_stall_s=round(self.client_stall_s, 3),
        )


def draw_jobs(n_jobs: int, rate: Optional[float], seed: int, grids=None, ks=None,
              obs_range=OBS_RANGE) -> List[JobSpec]:
    """Job shapes and Poisson arrival times (rate None: all arrive at t=0).
    The shapes depend on the seed only, so the same seed gives the same jobs
    at every rate."""
    grids = GRIDS if grids is None else grids
    ks = KS if ks is None else ks
    rng = np.random.Generator(np.random.Philox(np.random.SeedSequence([seed, 77])))
    g = rng.integers(0, len(grids), n_jobs)
    kk = rng.integers(0, len(ks), n_jobs)
    lo, hi = math.log(obs_range[0]), math.log(obs_range[1])
    nobs = np.exp(rng.uniform(lo, hi, n_jobs)).astype(np.int64)
    if rate is None:
        t = np.zeros(n_jobs)
    else:
        rt = np.random.Generator(np.random.Philox(np.random.SeedSequence([seed, 78])))
        t = np.cumsum(rt.exponential(1.0 / rate, n_jobs))
        t -= t[0]
    return [JobSpec(i, float(t[i]), int(g[i]), int(ks[kk[i]]), int(nobs[i])) for i in range(n_jobs)]


class Pools:
    """Device-resident inputs shared read-only by the jobs."""

    def __init__(self, torch, dev, seed: int, grids, max_obs: int):
        self.torch, self.dev, self.seed = torch, dev, seed
        self.grids = grids
        self.max_obs = int(max_obs)
        self.bg: Dict[Tuple[int, int], tuple] = {}
        self.obs: Dict[Tuple[int, int], dict] = {}

    def background(self, gi: int, k: int):
        key = (gi, k)
        if key not in self.bg:
            torch = self.torch
            nlat, nlon, nz = self.grids[gi]
            n = 1 + 3 * nz
            gen = torch.Generator(device=self.dev)
            gen.manual_seed(self.seed * 1009 + gi * 131 + k)
            xb = torch.randn((nlon, nlat, n, k), generator=gen, device=self.dev, dtype=torch.float64)
            xb -= xb.mean(dim=3, keepdim=True)
            axb = torch.randn((nlon, nlat, n), generator=gen, device=self.dev, dtype=torch.float64)
            lat, lon = synth.cell_centred_grid(nlat, nlon)
            self.bg[key] = (xb, axb, lat, lon)
        return self.bg[key]

    def obs_pool(self, k: int, nz: int) -> dict:
        key = (k, nz)
        if key not in self.obs:
            torch = self.torch
            rng = synth._rng(self.seed, 1000 + nz)
            nobs = self.max_obs
            n_st, per_st = 1000, 3 * nz
            n_sw = int(round(0.7 * nobs))
            n_un = max(nobs - n_sw - n_st * per_st, 0)
            sla, slo = synth._swath_points(rng, n_sw, 40, 50.0)
            tla, tlo = synth._uniform_sphere(rng, n_st)
            ula, ulo = synth._uniform_sphere(rng, n_un)
            la = np.concatenate([sla, tla, ula])
            lo = np.concatenate([slo, tlo, ulo])
            rows = np.concatenate([np.ones(n_sw, np.int64), np.full(n_st, per_st, np.int64),
                                   np.ones(n_un, np.int64)])
            perm = rng.permutation(la.size)
            la, lo, rows = la[perm], lo[perm], rows[perm]
            rp = np.zeros(la.size + 1, np.int64)
            np.cumsum(rows, out=rp[1:])
            nr = int(rp[-1])
            gen = torch.Generator(device=self.dev)
            gen.manual_seed(self.seed * 7919 + k * 17 + nz)
            sig = torch.where(torch.rand(nr, generator=gen, device=self.dev) < 1.0 / 3.0, 1.0, 2.0)
            yb = torch.randn((nr, k), generator=gen, device=self.dev, dtype=torch.float64)
            yb -= yb.mean(dim=1, keepdim=True)
            dep = torch.randn(nr, generator=gen, device=self.dev, dtype=torch.float64) * (sig * 1.4)
            rinv = (1.0 / (sig * sig)).to(torch.float64)
            vco = torch.ones(nr, device=self.dev, dtype=torch.float64)
            t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a.astype(dt))).to(self.dev)
            self.obs[key] = dict(lat=t(la, np.float64), lon=t(lo, np.float64), rp_host=rp,
                                 rp=t(rp, np.int64), yb=yb, dep=dep, rinv=rinv, vco=vco)
        return self.obs[key]

    def obs_slice(self, k: int, nz: int, n_rows: int, rng) -> Tuple[dict, list]:
        """A contiguous site range of the shuffled pool with about n_rows rows.
        Returns the pointer dict for the server and the tensors it points into."""
        torch = self.torch
        P = self.obs_pool(k, nz)
        rp = P["rp_host"]
        total = int(rp[-1])
        n_rows = min(int(n_rows), total)
        # first site s0 with room for n_rows rows after it
        s_max = int(np.searchsorted(rp, total - n_rows, side="right")) - 1
        s0 = int(rng.integers(0, max(s_max, 0) + 1))
        s1 = int(np.searchsorted(rp, rp[s0] + n_rows, side="left"))
        s1 = max(min(s1, rp.size - 1), s0)
        r0, r1 = int(rp[s0]), int(rp[s1])
        rel = (P["rp"][s0:s1 + 1] - r0).to(torch.int32)
        keep = [rel]
        d = dict(n_sites=s1 - s0, n_rows=r1 - r0,
                 site_lat=P["lat"][s0:].data_ptr(), site_lon=P["lon"][s0:].data_ptr(),
                 site_row_ptr=rel.data_ptr(), yb=P["yb"][r0:].data_ptr(),
                 dep=P["dep"][r0:].data_ptr(), rinv=P["rinv"][r0:].data_ptr(),
                 vcoord=P["vco"][r0:].data_ptr())
        return d, keep


def job_config(grid, k: int, precision: int):
    nlat, nlon, nz = grid
    cfg = synth.make_config("c2", k=k, precision=precision, nlat=nlat, nlon=nlon, nz=nz)
    cfg.local_dist = LOCAL_DIST_KM
    return cfg


def run(torch, jobs: List[JobSpec], pools: Pools, precision: int = PREC_BF16X3,
        hbm_budget: int = 96 << 30, max_active: int = 4, out_cap: int = 64 << 30,
        seed: int = 0, log=None) -> ServeResult:
    """Submit `jobs` at their arrival times; gather per-job latency."""
    from .engine import Server
    rng = np.random.Generator(np.random.Philox(np.random.SeedSequence([seed, 99])))
    res = ServeResult(n_jobs=len(jobs), rate=None)
    outstanding = {}
    out_bytes = 0

    def reap(srv, block=False):
        nonlocal out_bytes
        done = []
        for jid, rec in outstanding.items():
            st = srv.poll(jid)
            if st["state"] >= 2:
                done.append((jid, st))
        for jid, st in done:
            rec = outstanding.pop(jid)
            out_bytes -= rec["bytes"]
            if st["state"] != 2 or st["rc"] < 0:
                res.failed += 1
            arr = t0 + rec["spec"].arrival
            res.latency.append(st["t_end"] - arr)
            res.queue.append(st["t_start"] - arr)
            res.analyses += int(st["n_analyses"])
            res.makespan = max(res.makespan, st["t_end"] - t0)
            if log:
                sp = rec["spec"]
                log(f"job {sp.i} done: grid={pools.grids[sp.grid]} k={sp.k} obs={sp.n_obs} "
                    f"latency={st['t_end'] - arr:.2f}s queue={st['t_start'] - arr:.2f}s "
                    f"state={st['state']}")
        return len(done)

    with Server(hbm_budget, max_active) as srv:
        t0 = srv.now()
        for j in jobs:
            while srv.now() - t0 < j.arrival:
                if not reap(srv):
                    time.sleep(min(0.002, max(0.0, j.arrival - (srv.now() - t0))))
            grid = pools.grids[j.grid]
            xb, axb, glat, glon = pools.background(j.grid, j.k)
            need = xb.numel() * 8 + axb.numel() * 8
            if out_bytes + need > out_cap and outstanding:
                res.client_stalls += 1
                ts = srv.now()
                while out_bytes + need > out_cap and outstanding:
                    if not reap(srv):
                        time.sleep(0.002)
                res.client_stall_s += srv.now() - ts
            xa = torch.empty_like(xb)
            axa = torch.empty_like(axb)
            cyc, keep = [], [xa, axa]
            for _ in range(CYCLES):
                d, kk = pools.obs_slice(j.k, grid[2], j.n_obs, rng)
                cyc.append(d)
                keep += kk
            torch.cuda.current_stream().synchronize()
            cfg = job_config(grid, j.k, precision)
            jid = srv.submit(cfg, glat, glon, xb.data_ptr(), axb.data_ptr(), cyc, xa.data_ptr(),
                             axa.data_ptr())
            outstanding[jid] = dict(keep=keep, bytes=need, spec=j)
            out_bytes += need
        while outstanding:
            if not reap(srv):
                time.sleep(0.002)
    return res
