"""Independent pins for the oracle's analysis loop and the caller epilogue.

The C oracle's loop (``orc_letkf_compute``) and epilogue (``orc_epilogue``)
are checked here against

1. a second restatement in numpy, written from the reference text rather than
   from the oracle's C: ``letkf_compute_nooverlap`` (letkf_compute.F90:12-154),
   ``letkf_compute_nooverlap_vertical`` (letkf_compute.F90:386-554), the
   overlapped-cube ``letkf_compute`` (letkf_compute.F90:156-256) with
   ``xb_add2_xa`` (:332-382) and the ``cube_count`` division (:248-255),
   ``gen_yb_dis[_vertical]`` (letkf_yb.F90:3-238), ``vertical_localization``
   (letkf_yb.F90:240-260), ``schr2`` (letkf_yb.F90:262-286),
   ``latlon_dist_r8`` (pro_info.F90:148-161), ``letkf_core``
   (letkf_core.F90:30-225, with LAPACK's eigh in place of dsyevd, so the
   eigensolver is independent of the oracle's QL), ``xb_2_xa``
   (letkf_compute.F90:259-329) and ``ps2alnp`` (obs_make.F90:418-437 with
   the hybrid p(ps) of vert_coord_mod.F90:556-563).  Each pin is paired with
   a mutant of the restatement (the cube centre one row off, ps updated with
   the first vertical chunk instead of the last, polar obs kept, obs counted
   twice, rdiaginv kept in double, the ln p of a level instead of the layer
   midpoint, overlapped cubes not divided by their count, stepped by cube_n,
   or added without subtracting their mean) that the oracle must *disagree*
   with, so a slip of that kind in the oracle could not pass;
2. closed forms of the whole analysis: R^-1 = 0 leaves the state unchanged;
   the mean increment is Xb overline_w; one observation at a grid point with
   H = identity and rho = 1 gives the scalar ETKF update
   xa_mean = xb_mean + sb2 / (sb2 + so2) (y - xb_mean) and analysis variance
   sb2 so2 / (sb2 + so2); repeated cycles against a static truth reduce the
   analysis error (the toy-model check of x86/DA/LETKF/unit/main.f90:112-137);
3. closed forms of the epilogue (module_da_parallel.F90:219-261).
"""
import copy

import numpy as np
import pytest

# pro_info.F90:154-157: pi = atan(1.0)*4 and a = 6371.393 are real(4)
PI_F = float(np.float32(np.arctan(np.float32(1.0)) * np.float32(4.0)))
A_F = float(np.float32(6371.393))
# letkf_yb.F90:273-278: the literals 1.666667 and 0.6666667 are real(4)
C1 = float(np.float32(1.666667))
C2 = float(np.float32(0.6666667))


# ---------------------------------------------------------------------------
# restatement
# ---------------------------------------------------------------------------

def latlon_dist(x1, y1, x2, y2):
    """latlon_dist_r8 (pro_info.F90:148-161), elementwise."""
    v = (np.cos(x1 / 180 * PI_F) * np.cos(x2 / 180 * PI_F) * np.cos(y1 / 180 * PI_F - y2 / 180 * PI_F)
         + np.sin(x1 / 180 * PI_F) * np.sin(x2 / 180 * PI_F))
    return A_F * np.arccos(np.minimum(v, 1.0))


def schr2(ri):
    """Gaspari-Cohn (letkf_yb.F90:262-286), elementwise, same branch order."""
    ri = np.asarray(ri, dtype=np.float64)
    out = np.ones_like(ri)
    outer = (ri > 1.0) & (ri <= 2.0)
    inner = (ri > 0.0) & (ri <= 1.0)
    r = ri[outer]
    rj = r / 12.0 - 0.5
    rj = rj * r + 0.625
    rj = rj * r + C1
    out[outer] = (rj * r - 5.0) * r + 4.0 - C2 / r
    r = ri[inner]
    rk = r * r
    rj = (-0.25 * r + 0.5) * r + 0.625
    out[inner] = (rj * r - C1) * rk + 1.0
    out[ri > 2.0] = 0.0
    return out


def schr(ri):
    """Boxcar (letkf_yb.F90:287-302)."""
    return np.where(np.asarray(ri) > 2.0, 0.0, 1.0)


def gen_yb_dis(obs, clat, clon, cfg, mut):
    """gen_yb_dis: rows of every site within the GC support, in scan order
    (site, then its rows level-major), with the site's GC weight."""
    lat = obs.site_lat
    keep = np.abs(lat) <= cfg.obs_polar if "polar_kept" not in mut else np.ones(lat.size, bool)
    d = latlon_dist(lat, obs.site_lon, clat, clon)
    gc = schr2(d * 1.0 / cfg.local_dist) if cfg.gc_filter else schr(d * 1.0 / cfg.local_dist)
    rows, w = [], []
    rp = obs.site_row_ptr
    for s in np.flatnonzero(keep & (gc > 0)):
        rr = np.arange(rp[s], rp[s + 1])
        rows.append(rr)
        w.append(np.full(rr.size, gc[s]))
    if not rows:
        return np.zeros(0, np.int64), np.zeros(0)
    rows, w = np.concatenate(rows), np.concatenate(w)
    if "double_count" in mut:        # the one-rank duplication of every obs
        rows, w = np.concatenate([rows, rows]), np.concatenate([w, w])
    return rows, w


def letkf_core(Y, rdiaginv, dep, rloc, infl):
    """letkf_core.F90:58-225 in matrix form (Y is k x p)."""
    k = Y.shape[0]
    rr = np.sqrt(rdiaginv.astype(np.float64) * rloc)
    hx = Y * rr[None, :]
    work1 = hx @ hx.T
    lam, V = np.linalg.eigh(work1)
    lam = lam * 1.0 + (k - 1) / infl
    pa = (V / lam[None, :]) @ V.T
    hx = hx * rr[None, :]
    work2 = pa @ hx
    w = work2 @ dep
    T = (V * np.sqrt((k - 1) / lam)[None, :]) @ V.T
    return T + w[:, None], w


def update(xb_c, axb_c, trans):
    """xa(:, c) = trans^T xb(:, c) + axb(c) (letkf_compute.F90:290-302)."""
    return xb_c @ trans + axb_c[..., None]


def ps2alnp(ps, cfg):
    """ln p of the layer midpoints from the hybrid p(ps); the interface
    nz + 1, which ps2alnp reads unassigned, is taken from the coefficients."""
    hyai, hybi = np.asarray(cfg.hyai), np.asarray(cfg.hybi)
    ph = hyai * (cfg.p0 - cfg.ptop) + hybi * (ps - cfg.ptop) + cfg.ptop
    return np.log(0.5 * (ph[:-1] + ph[1:]))


def loop(inst, mut=()):
    """letkf_compute_nooverlap[_vertical], or letkf_compute when
    cfg.overlap > 0, over the instance.  Returns xa, axa (the analysed
    interior columns; halo entries 0), trans, overline_w and nobsl per
    analysis in the engine's order (lon cube major, then lat cube, then
    vertical chunk)."""
    cfg = inst.cfg
    k, h, cn = cfg.k, cfg.halo, cfg.cube_n
    ovl = getattr(cfg, "overlap", 0)
    # letkf_compute.F90:53,61 step cube_n; :197,205 step cube_n - overlap
    step = cn - ovl if ovl > 0 and "stride_cn" not in mut else cn
    rdi = inst.obs.rinv if "rinv_f64" in mut else inst.obs.rinv.astype(np.float32)
    xa = np.zeros_like(inst.xb)
    axa = np.zeros_like(inst.axb)
    count = np.zeros(inst.axb.shape[:2])
    out_t, out_w, out_n = [], [], []
    vs = cfg.var_start() - 1
    vl = cfg.var_lev()
    for i in range(1, cfg.nlon + 1, step):              # letkf_compute.F90:53 / :197
        si, ei = i, min(i + cn - 1, cfg.nlon)
        for j in range(1, cfg.nlat + 1, step):
            sj, ej = j, min(j + cn - 1, cfg.nlat)
            cj = (sj + ej) // 2 + (1 if "centre_off" in mut else 0)
            cj = min(cj, cfg.nlat)
            clat = float(inst.grid_lat[cj - 1])
            clon = float(inst.grid_lon[(si + ei) // 2 - 1])
            rows, gc = gen_yb_dis(inst.obs, clat, clon, cfg, mut)
            Y = inst.obs.yb[rows].T
            lonr = slice(si - 1 + h, ei + h)
            latr = slice(sj - 1 + h, ej + h)
            if not cfg.vertical:
                trans, w = letkf_core(Y, rdi[rows], inst.obs.dep[rows], gc, cfg.inflation_factor)
                out_t.append(trans), out_w.append(w), out_n.append(rows.size)
                xe = update(inst.xb[lonr, latr], inst.axb[lonr, latr], trans)
                if ovl > 0:
                    # xb_add2_xa (:332-382): add dist - mean and mean, count the cube
                    m = xe.sum(axis=-1) / k
                    xa[lonr, latr] += xe - (0.0 if "no_recentre" in mut else m[..., None])
                    axa[lonr, latr] += m
                    count[lonr, latr] += 1.0
                    continue
                m = xe.sum(axis=-1) / k
                xa[lonr, latr] = xe - m[..., None]
                axa[lonr, latr] = m
                continue
            # vertical loop: one selection per column, one core per chunk
            ocoord = inst.obs.vcoord[rows]
            if cfg.vcoord_mode == 0:
                alnp = np.arange(1, cfg.nz + 1, dtype=np.float64)
            else:
                ci_ = (si + ei) // 2
                ps = inst.axb[ci_ - 1 + h, (sj + ej) // 2 - 1 + h, cfg.ps_index]
                if "lnp_level" in mut:
                    hyai, hybi = np.asarray(cfg.hyai), np.asarray(cfg.hybi)
                    ph = hyai * (cfg.p0 - cfg.ptop) + hybi * (ps - cfg.ptop) + cfg.ptop
                    alnp = np.log(ph[:-1])
                else:
                    alnp = ps2alnp(ps, cfg)
            cz = cfg.cube_nz if cfg.cube_nz > 0 else cfg.nz
            xe_col = np.zeros(inst.xb[lonr, latr].shape)
            for kk in range(1, cfg.nz + 1, cz):
                sk, ek = kk, min(kk + cz - 1, cfg.nz)
                ck = (sk + ek) // 2
                f = schr2 if cfg.gc_filter_vertical else schr
                rl = gc * f(np.abs(alnp[ck - 1] - ocoord) / cfg.vert_dist)
                trans, w = letkf_core(Y, rdi[rows], inst.obs.dep[rows], rl, cfg.inflation_factor)
                out_t.append(trans), out_w.append(w), out_n.append(rows.size)
                for v in range(cfg.n_var):
                    if vl[v] == 1:
                        last = (sk == 1) if "ps_first_chunk" in mut else (ek == cfg.nz)
                        if not last:
                            continue
                        idx = [vs[v]]
                    else:
                        idx = list(range(vs[v] + sk - 1, vs[v] + ek))
                    xe_col[:, :, idx] = update(inst.xb[lonr, latr][:, :, idx],
                                               inst.axb[lonr, latr][:, :, idx], trans)
            m = xe_col.sum(axis=-1) / k                       # :539-552
            xa[lonr, latr] = xe_col - m[..., None]
            axa[lonr, latr] = m
    if ovl > 0 and "no_count" not in mut:
        # :248-255, interior columns only (the halo stays 0)
        inner = (slice(h, h + cfg.nlon), slice(h, h + cfg.nlat))
        xa[inner] /= count[inner][..., None, None]
        axa[inner] /= count[inner][..., None]
    return dict(xa=xa, axa=axa, trans=np.array(out_t), overline_w=np.array(out_w),
                nobsl=np.array(out_n))


def epilogue(inst, xa, axa, polar_lat, g1, g2):
    """run_da_parallel's letkf2dnn block (module_da_parallel.F90:219-261)."""
    cfg = inst.cfg
    h, k = cfg.halo, cfg.k
    out = np.zeros_like(inst.xb)
    sb = np.zeros((cfg.nlon, cfg.nlat, cfg.n))
    sa = np.zeros_like(sb)
    for jl in range(cfg.nlon):
        for il in range(cfg.nlat):
            c = (jl + h, il + h)
            xb, ab = inst.xb[c], inst.axb[c][:, None]
            if abs(float(inst.grid_lat[il])) <= polar_lat:
                o = g1 * (ab + xb) + g2 * (axa[c][:, None] + xa[c])
                spa = ((o - axa[c][:, None]) ** 2).sum(axis=1)
            else:
                o = xb + ab
                spa = (xb ** 2).sum(axis=1)
            out[c] = o
            sb[jl, il] = np.sqrt((xb ** 2).sum(axis=1) / float(np.float32(k)))
            sa[jl, il] = np.sqrt(spa / float(np.float32(k)))
    return out, sb, sa


# ---------------------------------------------------------------------------
# cases
# ---------------------------------------------------------------------------

def _horizontal(seed=1, **kw):
    from dida_letkf import synth
    args = dict(k=8, nlat=6, nlon=12, n_obs=90, local_dist=1500.0)
    args.update(kw)
    return synth.generate("c0", seed=seed, **args)


def _vertical(seed=2, lnp=False, **kw):
    from dida_letkf import synth
    args = dict(k=6, nlat=5, nlon=10, nz=5, n_obs=150)
    args.update(kw)
    inst = synth.generate("c1", seed=seed, **args)
    inst.cfg.local_dist = 2000.0
    inst.cfg.vert_dist = 1.5
    inst.cfg.cube_nz = 2          # chunks 1-2, 3-4, 5: ps rides on the last
    if lnp:
        from dida_letkf import config
        cfg = inst.cfg
        cfg.vcoord_mode = config.VCOORD_LNP
        cfg.hyai, cfg.hybi, cfg.p0, cfg.ptop = [list(c) if i < 2 else c for i, c in
                                                  enumerate(synth.hybrid_coefficients(cfg.nz))]
        cfg.ps_index = 0
        cfg.vert_dist = 0.4
        h = cfg.halo
        inst.axb[..., 0] += 1.0e5 + 500.0 * np.sin(np.arange(inst.axb.shape[0]))[:, None]
        rng = np.random.default_rng(seed)
        inst.obs.vcoord = np.log(rng.uniform(2.0e4, 1.0e5, inst.obs.n_rows))
    return inst


CASES = {
    "gc_cube1": lambda: _horizontal(1),
    "boxcar": lambda: _with(_horizontal(3), gc_filter=False),
    "cube2_clipped_halo_polar_inflation": lambda: _with(
        _horizontal(4, nlat=7, nlon=13, halo=1), cube_n=2, polar_discard=True, polar_lat=50.0,
        inflation_factor=1.3),
    "overlap_c4o2_halo": lambda: _with(_horizontal(9, nlat=9, nlon=14, halo=1), cube_n=4,
                                       overlap=2),
    "overlap_c3o1_clipped": lambda: _with(_horizontal(10, nlat=8, nlon=11), cube_n=3, overlap=1),
    "overlap_c5o3_inflation": lambda: _with(_horizontal(11, nlat=7, nlon=12, n_obs=300), cube_n=5,
                                            overlap=3, inflation_factor=1.2),
    "vertical_levels": lambda: _vertical(5),
    "vertical_lnp": lambda: _vertical(6, lnp=True),
}


def _with(inst, **kw):
    for key, v in kw.items():
        setattr(inst.cfg, key, v)
    return inst


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _compare(ours, ref):
    """Largest relative difference over xa, axa, trans and overline_w, or inf
    when the local obs counts differ."""
    if not np.array_equal(ours["nobsl"], ref["nobsl"]):
        return np.inf
    return max(_rel(ours["xa"], ref["xa"]), _rel(ours["axa"], ref["axa"]),
               _rel(ours["trans"], ref["trans"]), _rel(ours["overline_w"], ref["overline_w"]))


@pytest.mark.parametrize("case", sorted(CASES))
def test_loop_restatement_matches_oracle(orc, case):
    inst = CASES[case]()
    ref = orc.letkf_compute(inst, want_trans=True)
    ours = loop(inst)
    assert ours["nobsl"].max() > inst.cfg.k            # full-rank products somewhere
    err = _compare(ours, ref)
    print(case, "max rel diff", err)
    assert err < 1e-11


MUTANTS = [
    ("gc_cube1", "centre_off"),
    ("cube2_clipped_halo_polar_inflation", "centre_off"),
    ("cube2_clipped_halo_polar_inflation", "polar_kept"),
    ("gc_cube1", "double_count"),
    ("gc_cube1", "rinv_f64"),
    ("vertical_levels", "ps_first_chunk"),
    ("vertical_lnp", "ps_first_chunk"),
    ("vertical_lnp", "lnp_level"),
    ("overlap_c4o2_halo", "no_count"),
    ("overlap_c4o2_halo", "stride_cn"),
    ("overlap_c3o1_clipped", "no_recentre"),
    ("overlap_c5o3_inflation", "centre_off"),
]


@pytest.mark.parametrize("case,mut", MUTANTS)
def test_oracle_rejects_mutant(orc, case, mut):
    """Each slip, made in the restatement, moves it away from the oracle by
    far more than the pin's 1e-11: the pin would catch it in the oracle."""
    inst = CASES[case]()
    if mut == "rinv_f64":
        # sigma^2 values whose reciprocal is not a float: the real(4) cast shows
        inst.obs.rinv = 1.0 / np.linspace(0.31, 2.9, inst.obs.n_rows) ** 2
    ref = orc.letkf_compute(inst, want_trans=True)
    assert _compare(loop(inst), ref) < 1e-11
    err = _compare(loop(inst, mut=(mut,)), ref)
    print(case, mut, err)
    assert err > 1e-9


# ---------------------------------------------------------------------------
# closed forms of the whole analysis
# ---------------------------------------------------------------------------

def test_no_obs_weight_leaves_state(orc):
    """R^-1 = 0 (and rho = 1): A = 0, b = 0, trans = I, so xa = xb and
    axa = axb (xb is mean-zero)."""
    inst = _horizontal(7)
    inst.obs.rinv[:] = 0.0
    ref = orc.letkf_compute(inst, want_trans=True)
    assert ref["nobsl"].max() > 0
    assert np.abs(ref["xa"] - inst.xb).max() < 1e-13
    assert np.abs(ref["axa"] - inst.axb).max() < 1e-13
    assert np.abs(ref["trans"] - np.eye(inst.cfg.k)).max() < 1e-13


def test_overlap_no_obs_weight_leaves_state(orc):
    """Overlapped cubes with R^-1 = 0: every cube's trans is I, each adds
    dist = axb + xb (mean axb) once per covering cube, and the division by the
    cube count returns xb and axb."""
    inst = _with(_horizontal(12, nlat=9, nlon=13, halo=1), cube_n=4, overlap=2)
    inst.obs.rinv[:] = 0.0
    ref = orc.letkf_compute(inst, want_trans=True)
    h = inst.cfg.halo
    inner = (slice(h, h + inst.cfg.nlon), slice(h, h + inst.cfg.nlat))
    assert np.abs(ref["xa"][inner] - inst.xb[inner]).max() < 1e-13
    assert np.abs(ref["axa"][inner] - inst.axb[inner]).max() < 1e-13
    assert np.all(ref["xa"][:h] == 0) and np.all(ref["axa"][:, :h] == 0)


def test_mean_increment_is_xb_wbar(orc):
    """axa - axb = Xb overline_w per state element, because Yb and Xb are
    centred (T 1 = sqrt(rho) 1 and Xb 1 = 0)."""
    inst = _horizontal(8)
    inst.cfg.inflation_factor = 1.2
    ref = orc.letkf_compute(inst, want_trans=True)
    cfg = inst.cfg
    w = ref["overline_w"].reshape(cfg.nlon, cfg.nlat, cfg.k)
    inc = np.einsum("ijvk,ijk->ijv", inst.xb, w)
    assert np.abs(ref["axa"] - inst.axb - inc).max() < 1e-12 * max(1.0, np.abs(inc).max())


def test_single_obs_scalar_etkf(orc):
    """One observation at a grid point, H = identity, rho = 1 at distance 0:
    the analysed column follows the scalar Kalman update, mean
    xb + sb2/(sb2 + so2) (y - xb) and variance sb2 so2/(sb2 + so2)."""
    inst = _horizontal(9)
    cfg = inst.cfg
    jl, il = 5, 2                               # lon 5, lat 2 (0-based)
    xcol = inst.xb[jl, il, 0]                   # k perturbations of the 1 variable
    so2 = 0.25
    y = float(inst.axb[jl, il, 0]) + 0.7
    o = inst.obs
    o.site_lat = np.array([float(inst.grid_lat[il])])
    o.site_lon = np.array([float(inst.grid_lon[jl])])
    o.site_row_ptr = np.array([0, 1], np.int32)
    o.yb = xcol[None, :].copy()
    o.dep = np.array([y - float(inst.axb[jl, il, 0])])
    o.rinv = np.array([1.0 / so2])
    o.vcoord = np.ones(1)
    o.row_var = np.zeros(1, np.int32)
    ref = orc.letkf_compute(inst, want_trans=True)
    sb2 = float((xcol ** 2).sum() / (cfg.k - 1))
    gain = sb2 / (sb2 + so2)
    mean_a = float(inst.axb[jl, il, 0]) + gain * (y - float(inst.axb[jl, il, 0]))
    var_a = float((ref["xa"][jl, il, 0] ** 2).sum() / (cfg.k - 1))
    assert abs(ref["axa"][jl, il, 0] - mean_a) < 1e-13 * max(1.0, abs(mean_a))
    assert abs(var_a - sb2 * so2 / (sb2 + so2)) < 1e-13 * sb2
    # and the restatement agrees
    ours = loop(inst)
    assert abs(ours["axa"][jl, il, 0] - mean_a) < 1e-13 * max(1.0, abs(mean_a))


def test_cycling_reduces_error(orc):
    """Repeated analyses of a static truth, perfectly observed at grid
    points, each cycle's analysis the next background (the obs re-centred on
    the cycle's background): the analysis error falls every cycle, as in the
    reference's toy-model test (x86/DA/LETKF/unit/main.f90:112-137)."""
    inst = _horizontal(10, nlat=6, nlon=12)
    cfg = inst.cfg
    rng = np.random.default_rng(3)
    sites = rng.choice(cfg.nlat * cfg.nlon, size=40, replace=False)
    jl, il = sites // cfg.nlat, sites % cfg.nlat
    so = 0.3                                    # assumed error; the obs are perfect
    y = inst.truth[jl, il, 0].copy()
    o = inst.obs
    o.site_lat = inst.grid_lat[il].astype(np.float64)
    o.site_lon = inst.grid_lon[jl].astype(np.float64)
    o.site_row_ptr = np.arange(sites.size + 1, dtype=np.int32)
    o.rinv = np.full(sites.size, 1.0 / so ** 2)
    o.vcoord = np.ones(sites.size)
    o.row_var = np.zeros(sites.size, np.int32)
    cfg.inflation_factor = 1.1
    err = []
    cur = copy.copy(inst)
    for c in range(6):
        o.yb = cur.xb[jl, il, 0].copy()
        o.dep = y - cur.axb[jl, il, 0]
        cur.obs = o
        ref = orc.letkf_compute(cur, want_trans=False)
        err.append(float(np.sqrt(np.mean((ref["axa"][..., 0] - inst.truth[..., 0]) ** 2))))
        cur = copy.copy(cur)
        cur.xb, cur.axb = ref["xa"], ref["axa"]
    e0 = float(np.sqrt(np.mean((inst.axb[..., 0] - inst.truth[..., 0]) ** 2)))
    print("background error", e0, "analysis errors", err)
    assert err[0] < e0
    assert all(b < a for a, b in zip(err, err[1:]))
    assert err[-1] < 0.7 * e0


# ---------------------------------------------------------------------------
# epilogue
# ---------------------------------------------------------------------------

def test_epilogue_restatement_and_closed_forms(orc):
    inst = _horizontal(11, halo=1)
    cfg = inst.cfg
    ref = orc.letkf_compute(inst)
    h = cfg.halo
    for pl, g1, g2 in ((90.0, 0.0, 1.0), (40.0, 0.3, 0.7), (-1.0, 0.5, 0.5)):
        xo, sb, sa = orc.epilogue(inst, ref["xa"], ref["axa"], polar_lat=pl, gamma1=g1,
                                  gamma2=g2)
        xo_n, sb_n, sa_n = epilogue(inst, ref["xa"], ref["axa"], pl, g1, g2)
        assert _rel(xo, xo_n) < 1e-14 and _rel(sb, sb_n) < 1e-14 and _rel(sa, sa_n) < 1e-14
        inner = (slice(h, -h), slice(h, -h))
        # halo columns are zero
        assert np.all(xo[:h] == 0) and np.all(xo[-h:] == 0)
        assert np.all(xo[:, :h] == 0) and np.all(xo[:, -h:] == 0)
        lat_ok = np.abs(inst.grid_lat.astype(np.float64)) <= pl
        full_a = ref["axa"][inner][..., None] + ref["xa"][inner]
        full_b = inst.axb[inner][..., None] + inst.xb[inner]
        if g1 == 0.0 and g2 == 1.0:
            # the default blend is the analysis itself, mean + perturbation
            assert np.abs(xo[inner] - full_a).max() < 1e-13 * np.abs(full_a).max()
        # above the polar latitude the background comes back unchanged
        if (~lat_ok).any():
            d = np.abs(xo[inner][:, ~lat_ok] - full_b[:, ~lat_ok]).max()
            assert d <= 1e-13 * np.abs(full_b).max()
        # spread_b is the RMS of the background perturbations (divisor ens)
        sbx = np.sqrt((inst.xb[inner] ** 2).sum(axis=-1) / cfg.k)
        assert np.abs(sb - sbx).max() < 1e-14 * sbx.max()
