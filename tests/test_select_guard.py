"""The selection kernels decide most candidate sites with a cheap cos(angle)
(the cosine of the longitude difference expanded with precomputed cos/sin of
each longitude) and run the reference's exact latlon_dist_r8 arithmetic only
inside a guard band of 1e-12 around the support threshold.  That is exact
(the same sites as the reference) iff the cheap and exact cos(angle) differ
by far less than the band: checked here in FP64 over random and adversarial
points (same formulas as letkf_device.h, latlon_dist_pre)."""
import numpy as np

PI_F32 = 3.1415927410125732
BAND = 1e-12


def _rad(x):
    return x / 180.0 * PI_F32


def _exact_v(la1, lo1, la2, lo2):
    c1, s1 = np.cos(_rad(la1)), np.sin(_rad(la1))
    c2, s2 = np.cos(_rad(la2)), np.sin(_rad(la2))
    return c1 * c2 * np.cos(_rad(lo1) - _rad(lo2)) + s1 * s2


def _fast_v(la1, lo1, la2, lo2):
    c1, s1 = np.cos(_rad(la1)), np.sin(_rad(la1))
    c2, s2 = np.cos(_rad(la2)), np.sin(_rad(la2))
    l1, l2 = _rad(lo1), _rad(lo2)
    return c1 * c2 * (np.cos(l1) * np.cos(l2) + np.sin(l1) * np.sin(l2)) + s1 * s2


def test_cheap_cos_angle_is_within_the_band():
    rng = np.random.default_rng(0)
    n = 2_000_000
    la1, la2 = rng.uniform(-90, 90, n), rng.uniform(-90, 90, n)
    lo1, lo2 = rng.uniform(0, 360, n), rng.uniform(0, 360, n)
    # and pairs close together (the support of a local analysis)
    la3 = la1 + rng.normal(0, 3, n)
    lo3 = lo1 + rng.normal(0, 3, n)
    for a, b, c, d in ((la1, lo1, la2, lo2), (la1, lo1, la3, lo3)):
        err = np.abs(_fast_v(a, b, c, d) - _exact_v(a, b, c, d))
        assert err.max() < 1e-14, err.max()          # 100x below the band
    # a NaN position gives a NaN v, which neither side of the band holds
    v = _fast_v(np.nan, 0.0, 10.0, 20.0)
    assert not (v < -1.0 - BAND) and not (v > 1.0)
