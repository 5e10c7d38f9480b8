"""Pins the CPU oracle against closed forms, textbook formulas, invariants and
independent library routines (the reference ships no golden vectors for this
path).  Each test is chosen so that a plausible mistake in the oracle (a
dropped term, a wrong sign or index, a transposed operand, a missing float
cast) fails it.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

# ----------------------------------------------------------------------------
# Gaspari-Cohn (letkf_yb.F90:262-302)
# ----------------------------------------------------------------------------


def gc_textbook(r):
    """Gaspari & Cohn (1999) eq. 4.10 in expanded power form, exact constants."""
    r = abs(r)
    if r <= 1.0:
        return -0.25 * r**5 + 0.5 * r**4 + 0.625 * r**3 - 5.0 / 3.0 * r**2 + 1.0
    if r <= 2.0:
        return (r**5 / 12.0 - 0.5 * r**4 + 0.625 * r**3 + 5.0 / 3.0 * r**2 - 5.0 * r + 4.0
                - 2.0 / (3.0 * r))
    return 0.0


def test_schr2_special_values(orc):
    assert orc.schr2(0.0) == 1.0
    assert orc.schr2(-0.5) == 1.0          # ri <= 0 branch
    assert orc.schr2(2.0000001) == 0.0
    assert orc.schr2(50.0) == 0.0
    assert orc.schr(2.0) == 1.0 and orc.schr(2.0 + 1e-12) == 0.0 and orc.schr(0.3) == 1.0


def test_schr2_matches_textbook_gc(orc):
    rs = np.linspace(0.0, 2.0, 4001)[1:]
    err = max(abs(orc.schr2(r) - gc_textbook(r)) for r in rs)
    # only the real(4) literals 1.666667 / 0.6666667 separate the two
    assert err < 3e-6
    assert err > 1e-8      # ... and they do separate them: the literals are float


def test_schr2_float_literal_residual_at_support_edge(orc):
    # exact GC is 0 at r=2; the reference's real(4) literals leave
    # 4 - 2*5 + ... - 0.6666667f/2 = 1.26e-6 (computed in exact rationals)
    from fractions import Fraction as F
    import struct
    f32 = lambda x: F(struct.unpack('f', struct.pack('f', x))[0])
    r = F(2)
    rj = r / 12 - F(1, 2)
    rj = rj * r + F(5, 8)
    rj = rj * r + f32(1.666667)
    exact = (rj * r - 5) * r + 4 - f32(0.6666667) / r
    assert abs(orc.schr2(2.0) - float(exact)) < 1e-15
    assert 1.2e-6 < orc.schr2(2.0) < 1.3e-6


def test_schr2_positive_on_whole_support(orc):
    # the selection test gc > 0 (letkf_yb.F90:98) is therefore ri <= 2
    rs = np.linspace(1.0, 2.0, 100001)
    assert min(orc.schr2(r) for r in rs) > 0.0


def test_schr2_continuity(orc):
    eps = 1e-9
    assert abs(orc.schr2(1.0) - orc.schr2(1.0 + eps)) < 1e-6
    assert abs(orc.schr2(1.0) - gc_textbook(1.0)) < 1e-6
    assert abs(gc_textbook(1.0) - 5.0 / 24.0) < 1e-15


# ----------------------------------------------------------------------------
# latlon_dist_r8 (pro_info.F90:148-161)
# ----------------------------------------------------------------------------

def test_float_constants(orc):
    lib = orc.lib()
    assert lib.orc_pi_f32() == float(np.float32(np.pi))
    assert lib.orc_pi_f32() == 3.1415927410125732
    assert lib.orc_earth_radius_km() == float(np.float32(6371.393))


def test_latlon_closed_forms(orc):
    a = float(np.float32(6371.393))
    pif = float(np.float32(np.pi))
    # quarter circumference: cos(pi_f/2) = -4.37e-8 -> acos = pi/2 + 4.37e-8 = pi_f/2
    assert abs(orc.latlon_dist(0.0, 0.0, 0.0, 90.0) - a * pif / 2) < 1e-8
    # pole to a point on the equator: v = cos(pi_f/2) cos(17 pi_f/180) with
    # cos(pi_f/2) = -(pi_f - pi)/2 to first order
    v = -(pif - math.pi) / 2 * math.cos(17.0 / 180.0 * pif)
    assert abs(orc.latlon_dist(0.0, 0.0, 90.0, 17.0) - a * (math.pi / 2 - v)) < 1e-8
    # antipodes along a meridian: pole to pole
    assert abs(orc.latlon_dist(-90.0, 0.0, 90.0, 0.0) - a * math.pi) < 1e-3
    # zero distance (acos(min(.,1)) -> small but not necessarily 0)
    assert orc.latlon_dist(33.3, 120.1, 33.3, 120.1) < 1e-3


def test_latlon_symmetry_and_haversine(orc):
    rng = np.random.default_rng(0)
    a = float(np.float32(6371.393))
    scale = float(np.float32(np.pi)) / np.pi
    for _ in range(400):
        la1, la2 = rng.uniform(-89, 89, 2)
        lo1, lo2 = rng.uniform(0, 360, 2)
        d = orc.latlon_dist(la1, lo1, la2, lo2)
        assert d == orc.latlon_dist(la2, lo2, la1, lo1)
        # haversine with degrees converted by pi_f (same sphere, stable formula)
        p1, p2 = np.deg2rad(la1) * scale, np.deg2rad(la2) * scale
        dl = np.deg2rad(lo1 - lo2) * scale
        h = np.sin((p2 - p1) / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(dl / 2) ** 2
        dh = 2 * a * np.arcsin(np.sqrt(h))
        assert abs(d - dh) < 1e-9 * d + 2e-4


# ----------------------------------------------------------------------------
# eigensolvers
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 8, 20, 33, 64])
@pytest.mark.parametrize("method", ["ql", "jacobi"])
def test_syev_vs_lapack(orc, n, method):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + 0.1 * np.eye(n)
    w, v, info = orc.syev(A, method)
    assert info == 0
    wl = sla.eigh(A, eigvals_only=True)
    assert np.allclose(w, wl, rtol=0, atol=1e-12 * max(1.0, abs(wl).max()))
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-13
    assert np.abs(A @ v - v * w).max() < 1e-12 * np.abs(A).max()


def test_syev_degenerate_cluster(orc):
    # rank-3 PSD 20x20: 17-fold zero eigenvalue (the p < k case of the LETKF)
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((20, 3))
    A = Y @ Y.T
    for m in ("ql", "jacobi"):
        w, v, info = orc.syev(A, m)
        assert info == 0
        assert np.abs(v.T @ v - np.eye(20)).max() < 1e-13
        assert np.abs((v * w) @ v.T - A).max() < 1e-12 * np.abs(A).max()


# ----------------------------------------------------------------------------
# letkf_core (letkf_core.F90:14-233)
# ----------------------------------------------------------------------------

def letkf_numpy(Y, rinv32, dep, rloc, rho=1.0):
    """Independent algebra: Pa = M^-1 by solve, T = sqrt(k-1) M^-1/2 by eigh,
    wbar = Pa (Y R d)."""
    k = Y.shape[0]
    r = rinv32.astype(np.float64) * rloc
    M = (Y * r) @ Y.T + (k - 1) / rho * np.eye(k)
    wbar = np.linalg.solve(M, (Y * r) @ dep)
    lam, V = np.linalg.eigh(M)
    T = (V * np.sqrt((k - 1) / lam)) @ V.T
    return T + wbar[:, None], wbar, M, T


def _case(k, p, seed):
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((k, p))
    Y -= Y.mean(axis=0, keepdims=True)          # centred like letkf_obs.F90:156-183
    rinv = (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32)
    dep = rng.standard_normal(p)
    rloc = rng.uniform(0.0, 1.0, p)
    return Y, rinv, dep, rloc


@pytest.mark.parametrize("k,p", [(2, 1), (3, 17), (8, 5), (20, 49), (32, 300), (64, 500)])
def test_letkf_core_vs_numpy(orc, k, p):
    Y, rinv, dep, rloc = _case(k, p, 10 * k + p)
    trans, w, info = orc.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.1)
    assert info == 0
    tr_np, w_np, M, T = letkf_numpy(Y, rinv, dep, rloc, rho=1.1)
    assert np.abs(w - w_np).max() < 1e-12 * max(1, np.abs(w_np).max())
    assert np.abs(trans - tr_np).max() < 1e-12


def test_letkf_core_invariants(orc):
    k, p = 20, 73
    Y, rinv, dep, rloc = _case(k, p, 3)
    trans, w, _ = orc.letkf_core(Y, rinv, dep, rloc)
    T = trans - w[:, None]
    r = rinv.astype(np.float64) * rloc
    M = (Y * r) @ Y.T + (k - 1) * np.eye(k)
    Pa = np.linalg.inv(M)
    assert np.abs(T - T.T).max() < 1e-13
    assert np.abs(T @ T - (k - 1) * Pa).max() < 1e-12
    # Y is centred: M 1 = (k-1) 1  =>  T 1 = 1
    assert np.abs(T @ np.ones(k) - 1.0).max() < 1e-12


def test_letkf_core_p0_fast_path_values(orc):
    for k in (2, 20, 64):
        trans, w, info = orc.letkf_core(np.zeros((k, 0)), np.zeros(0, np.float32), np.zeros(0),
                                        np.zeros(0), inflation_factor=1.3)
        assert np.abs(w).max() == 0.0
        assert np.abs(trans - math.sqrt(1.3) * np.eye(k)).max() < 1e-14


def test_letkf_core_single_obs_sherman_morrison(orc):
    k = 16
    rng = np.random.default_rng(7)
    y = rng.standard_normal((k, 1))
    y -= y.mean()
    rinv = np.array([0.25], np.float32)
    d = np.array([1.7])
    rl = np.array([0.6])
    trans, w, _ = orc.letkf_core(y, rinv, d, rl)
    r = 0.25 * 0.6
    yy = float((y * y).sum())
    w_cf = (r * d[0] / ((k - 1) + r * yy)) * y[:, 0]
    s = r * yy / (k - 1)
    yh = y[:, 0] / math.sqrt(yy)
    T_cf = np.eye(k) + (1.0 / math.sqrt(1.0 + s) - 1.0) * np.outer(yh, yh)
    assert np.abs(w - w_cf).max() < 1e-14
    assert np.abs(trans - (T_cf + w_cf[:, None])).max() < 1e-13


def test_letkf_core_rdiaginv_is_real4(orc):
    # R^-1 = 1/3 is not float-exact: the oracle must use float(1/3) like
    # rdiaginv(real(4)) (letkf_yb.F90:10,108)
    k, p = 8, 12
    Y, _, dep, rloc = _case(k, p, 11)
    r64 = np.full(p, 1.0 / 3.0)
    trans, w, _ = orc.letkf_core(Y, r64.astype(np.float32), dep, rloc)
    tr32, w32, _, _ = letkf_numpy(Y, r64.astype(np.float32), dep, rloc)
    k_ = Y.shape[0]
    r = r64 * rloc
    M = (Y * r) @ Y.T + (k_ - 1) * np.eye(k_)
    w64 = np.linalg.solve(M, (Y * r) @ dep)
    assert np.abs(w - w32).max() < 1e-14
    assert np.abs(w - w64).max() > 1e-10


def test_letkf_core_jacobi_equals_ql(orc):
    Y, rinv, dep, rloc = _case(64, 900, 21)
    t1, w1, _ = orc.letkf_core(Y, rinv, dep, rloc)
    t2, w2, _ = orc.letkf_core(Y, rinv, dep, rloc, jacobi=True)
    assert np.abs(t1 - t2).max() < 1e-12
    assert np.abs(w1 - w2).max() < 1e-12


def test_letkf_core_rank_deficient(orc):
    # p < k-1: (k-1-p)-fold degenerate eigenvalue k-1 of M
    Y, rinv, dep, rloc = _case(32, 5, 9)
    t1, w1, _ = orc.letkf_core(Y, rinv, dep, rloc)
    t2, _, M, _ = letkf_numpy(Y, rinv, dep, rloc)
    assert np.abs(t1 - t2).max() < 1e-12


def test_letkf_core_mpal_guard(orc):
    # a non-PSD obs-space product (as reduced-precision products can give)
    # drives eigenvalues of A + (k-1)I below k-1; the guard
    # (letkf_core_mpal.F90:113-118) lifts the smallest to exactly k-1
    k = 6
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    lamA = np.array([-2.0, -0.5, 0.0, 1.0, 3.0, 7.0])
    A = (Q * lamA) @ Q.T
    b = rng.standard_normal(k)
    tg, wg, _ = orc.letkf_transform(A, b, guard=True)
    lam = lamA + (k - 1)
    lam2 = lam - (lam[0] - (k - 1))
    T_cf = (Q * np.sqrt((k - 1) / lam2)) @ Q.T
    w_cf = (Q / lam2) @ (Q.T @ b)
    assert np.abs(wg - w_cf).max() < 1e-13
    assert np.abs(tg - (T_cf + w_cf[:, None])).max() < 1e-13
    # PSD input: the guard is inactive and matches letkf_core
    Y, rinv, dep, rloc = _case(k, 30, 8)
    r = rinv.astype(np.float64) * rloc
    t1, w1, _ = orc.letkf_core(Y, rinv, dep, rloc)
    t2, w2, _ = orc.letkf_transform((Y * r) @ Y.T, (Y * r) @ dep, guard=True)
    assert np.abs(t1 - t2).max() < 1e-13 and np.abs(w1 - w2).max() < 1e-13


def test_letkf_transform_equals_core(orc):
    Y, rinv, dep, rloc = _case(32, 211, 12)
    r = rinv.astype(np.float64) * rloc
    t1, w1, _ = orc.letkf_core(Y, rinv, dep, rloc, inflation_factor=1.2)
    t2, w2, _ = orc.letkf_transform((Y * r) @ Y.T, (Y * r) @ dep, inflation_factor=1.2)
    assert np.abs(t1 - t2).max() < 1e-13 and np.abs(w1 - w2).max() < 1e-13


# ----------------------------------------------------------------------------
# xb_2_xa (letkf_compute.F90:259-329)
# ----------------------------------------------------------------------------

def test_xb_2_xa(orc):
    k, m = 20, 37
    rng = np.random.default_rng(1)
    xb = rng.standard_normal((m, k))
    xb -= xb.mean(axis=1, keepdims=True)
    axb = rng.standard_normal(m)
    Y, rinv, dep, rloc = _case(k, 40, 2)
    trans, w, _ = orc.letkf_core(Y, rinv, dep, rloc)
    xa, axa = orc.xb_2_xa(xb, axb, trans)
    full = xb @ trans + axb[:, None]                  # xa(e) = sum_f trans(f,e) xb(f)
    assert np.abs(axa - full.mean(axis=1)).max() < 1e-13
    assert np.abs(xa - (full - full.mean(axis=1, keepdims=True))).max() < 1e-13
    # T 1 = 1 for centred Yb  =>  mean analysis = axb + Xb wbar
    assert np.abs(axa - (axb + xb @ w)).max() < 1e-12
    assert np.abs(xa.sum(axis=1)).max() < 1e-12
