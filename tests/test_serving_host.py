"""Host logic of the serving workload: job draws are seeded, inside the
stated ranges, and arrivals form a Poisson stream of the stated rate."""
import numpy as np

from dida_letkf import serving


def test_draws_are_seeded_and_in_range():
    a = serving.draw_jobs(400, 2.0, seed=5)
    b = serving.draw_jobs(400, 2.0, seed=5)
    assert [(j.grid, j.k, j.n_obs, j.arrival) for j in a] == [(j.grid, j.k, j.n_obs, j.arrival) for j in b]
    assert {j.k for j in a} == set(serving.KS)
    assert {j.grid for j in a} == set(range(len(serving.GRIDS)))
    n = np.array([j.n_obs for j in a])
    assert n.min() >= serving.OBS_RANGE[0] - 1 and n.max() <= serving.OBS_RANGE[1]
    # log-uniform: the median is near the geometric mean of the range
    gm = np.sqrt(serving.OBS_RANGE[0] * serving.OBS_RANGE[1])
    assert 0.5 * gm < np.median(n) < 2.0 * gm


def test_poisson_arrivals():
    rate = 3.0
    jobs = serving.draw_jobs(4000, rate, seed=9)
    t = np.array([j.arrival for j in jobs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() * rate - 1.0) < 0.05          # mean gap 1/rate
    assert abs(gaps.std() * rate - 1.0) < 0.08           # exponential: sd = mean
    burst = serving.draw_jobs(10, None, seed=9)
    assert all(j.arrival == 0.0 for j in burst)


def test_job_config_matches_workload():
    cfg = serving.job_config(serving.GRIDS[1], 64, 2)
    assert (cfg.nlat, cfg.nlon, cfg.nz, cfg.k, cfg.n) == (180, 360, 32, 64, 97)
    assert cfg.local_dist == serving.LOCAL_DIST_KM and cfg.vert_dist < 0


def test_same_seed_same_shapes_at_every_rate():
    a = serving.draw_jobs(50, None, seed=3)
    b = serving.draw_jobs(50, 0.7, seed=3)
    assert [(j.grid, j.k, j.n_obs) for j in a] == [(j.grid, j.k, j.n_obs) for j in b]
