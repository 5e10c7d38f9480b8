/*
 * letkf_oracle.c -- CPU restatement of the reference LETKF analysis path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never calls into it.
 *
 * Every routine cites the reference routine it restates (paths relative to
 * the reference repository root).  Arithmetic is plain IEEE double with
 * -ffp-contract=off, in the operation order of the Fortran source, including
 * the reference's single-precision literals:
 *   - latlon_dist_r8  x86/DA/utils/pro_info.F90:148-161
 *       pi = atan(1.0)*4 is evaluated in real(4)  -> 3.1415927410125732
 *       a  = 6371.393 is a real(4) literal        -> 6371.39306640625 km
 *   - schr2 / schr    x86/DA/LETKF/src/letkf_yb.F90:262-302
 *       1.666667 and 0.6666667 are real(4) literals promoted to real(8)
 *   - gen_yb_dis      x86/DA/LETKF/src/letkf_yb.F90:3-150 (rdiaginv is real(4))
 *   - letkf_core      x86/DA/LETKF/src/letkf_core.F90:14-233
 *   - letkf_core_mpal eigenvalue guard, x86/DA/LETKF/src/letkf_core_mpal.F90:73-76,113-118
 *   - xb_2_xa         x86/DA/LETKF/src/letkf_compute.F90:259-329
 *   - letkf_compute_nooverlap          letkf_compute.F90:12-154
 *   - letkf_compute (overlapped cubes) letkf_compute.F90:156-256, with
 *     xb_add2_xa :332-382 and the cube_count division :248-255
 *   - letkf_compute_nooverlap_vertical letkf_compute.F90:386-554
 *   - vertical_localization            letkf_yb.F90:240-260
 *   - run_da_parallel epilogue         x86/DA/EnDNN/src/module_da_parallel.F90:186-261
 *
 * The reference computes the k x k eigendecomposition with LAPACK dsyevd.
 * LAPACK is a third-party dependency that is not part of the reference tree,
 * so the oracle carries two eigensolvers of its own, each written from the
 * published method:
 *   - Householder tridiagonalisation + implicit QL with Wilkinson shift
 *     (Bowdler, Martin, Reinsch & Wilkinson 1968; Golub & Van Loan, 4th ed.,
 *     Alg. 8.3.1 and Sec. 8.3.5), the default;
 *   - cyclic Jacobi (Golub & Van Loan Alg. 8.5.3), used to cross-check.
 * The LETKF output (trans, overline_w) is invariant to the eigenvector basis,
 * so parity is defined on trans and overline_w, not on the eigenvectors.
 *
 * Parity notes (deliberate, documented differences from the reference):
 *   - The reference duplicates every obs when run on one rank in longitude
 *     (letkf_init.F90:146-152 with letkf_utils.F90:2-16).  The oracle counts
 *     each obs once (the intended single-count semantics).
 *   - MAX_OBS truncation (letkf_yb.F90:124) is optional: max_obs <= 0 means
 *     no cap.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <pthread.h>
#include <float.h>

#include "letkf_oracle.h"

/* ------------------------------------------------------------------------ */
/* Localisation and distance                                                 */
/* ------------------------------------------------------------------------ */

/* letkf_yb.F90:262-286 */
double orc_schr2(double ri)
{
    const double c_1666667 = (double)1.666667f;
    const double c_0666667 = (double)0.6666667f;
    double rci, rj, rk;
    if (ri > 2.0) {
        rci = 0.0;
    } else if (ri > 1.0) {
        rj = ri / 12.0 - 0.5;
        rj = rj * ri + 0.625;
        rj = rj * ri + c_1666667;
        rci = (rj * ri - 5.0) * ri + 4.0 - c_0666667 / ri;
    } else if (ri > 0.0) {
        rk = ri * ri;
        rj = (-0.25 * ri + 0.5) * ri + 0.625;
        rci = (rj * ri - c_1666667) * rk + 1.0;
    } else {
        rci = 1.0;
    }
    return rci;
}

/* letkf_yb.F90:287-302 (boxcar) */
double orc_schr(double ri)
{
    return (ri > 2.0) ? 0.0 : 1.0;
}

/* pro_info.F90:148-161 */
double orc_pi_f32(void)
{
    float q = atanf(1.0f);
    float p = q * 4.0f;
    return (double)p;
}

double orc_earth_radius_km(void)
{
    return (double)6371.393f;
}

double orc_latlon_dist(double x1, double y1, double x2, double y2)
{
    const double pi = orc_pi_f32();
    const double a = orc_earth_radius_km();
    double c1 = cos(x1 / 180.0 * pi);
    double c2 = cos(x2 / 180.0 * pi);
    double c3 = cos(y1 / 180.0 * pi - y2 / 180.0 * pi);
    double s1 = sin(x1 / 180.0 * pi);
    double s2 = sin(x2 / 180.0 * pi);
    double v = c1 * c2 * c3 + s1 * s2;
    if (v > 1.0) v = 1.0;          /* min(v, 1.0) */
    return a * acos(v);
}

/* ------------------------------------------------------------------------ */
/* Local observation selection                                               */
/* ------------------------------------------------------------------------ */

/*
 * gen_yb_dis (letkf_yb.F90:3-150) and gen_yb_dis_vertical (:154-238) over a
 * single global obs list, scanned in obs order; every row (level, var) of a
 * site with gc > 0 is appended in site-row order.  rdiaginv is real(4)
 * (letkf_yb.F90:10,108).  Returns nobsl.
 */
int orc_gen_yb_dis(const orc_obs *obs, int k, double center_lat, double center_lon,
                   double local_dist, int gc_filter, double obs_polar, int max_obs,
                   int cap, double *yb, double *dis, float *rdiaginv, double *rloc,
                   double *vco)
{
    int nobsl = 0;
    for (int s = 0; s < obs->n_sites; ++s) {
        double la = obs->site_lat[s], lo = obs->site_lon[s];
        if (!(fabs(la) <= obs_polar)) continue;
        double d = orc_latlon_dist(la, lo, center_lat, center_lon);
        double ri = d * 1.0 / local_dist;
        double gc = gc_filter ? orc_schr2(ri) : orc_schr(ri);
        if (!(gc > 0.0)) continue;
        int r0 = obs->site_row_ptr ? obs->site_row_ptr[s] : s;
        int r1 = obs->site_row_ptr ? obs->site_row_ptr[s + 1] : s + 1;
        for (int r = r0; r < r1; ++r) {
            if (nobsl >= cap) return -1;
            memcpy(yb + (size_t)nobsl * k, obs->yb + (size_t)r * k, sizeof(double) * k);
            dis[nobsl] = obs->dep[r];
            rdiaginv[nobsl] = (float)obs->rinv[r];
            rloc[nobsl] = gc;
            if (vco) vco[nobsl] = obs->vcoord ? obs->vcoord[r] : 0.0;
            ++nobsl;
            if (max_obs > 0 && nobsl >= max_obs) return nobsl;
        }
    }
    return nobsl;
}

/* vertical_localization, letkf_yb.F90:240-260 */
void orc_vertical_localization(const double *rloc_in, double *rloc, int nobsl,
                               double lnp, const double *obs_lnp, double vert_dist,
                               int gc_filter_vertical)
{
    for (int i = 0; i < nobsl; ++i) {
        double ri = fabs(lnp - obs_lnp[i]) / vert_dist;
        rloc[i] = rloc_in[i] * (gc_filter_vertical ? orc_schr2(ri) : orc_schr(ri));
    }
}

/* ------------------------------------------------------------------------ */
/* Symmetric eigensolvers (column-major n x n, ascending eigenvalues)        */
/* ------------------------------------------------------------------------ */

#define IDX(i, j, ld) ((size_t)(j) * (ld) + (i))

/*
 * Householder reduction of the symmetric matrix a (lower triangle used) to
 * tridiagonal d (diag), e (e[i] couples i and i+1), and the explicit
 * orthogonal q with a = q * T * q^T.  Golub & Van Loan Alg. 8.3.1.
 * work: 3n doubles.
 */
static void orc_tridiag(int n, double *a, double *d, double *e, double *q, double *work)
{
    double *v = work, *p = work + n, *w = work + 2 * n;
    double *tau = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *vs = (double *)malloc(sizeof(double) * (size_t)n * n);
    /* symmetrise from the lower triangle */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) a[IDX(i, j, n)] = a[IDX(j, i, n)];
    for (int j = 0; j + 2 < n; ++j) {
        int m = n - j - 1;                 /* length of x = a[j+1:n, j] */
        double sig = 0.0;
        for (int i = 0; i < m; ++i) { v[i] = a[IDX(j + 1 + i, j, n)]; sig += v[i] * v[i]; }
        double x0 = v[0];
        double nrm = sqrt(sig);
        tau[j] = 0.0;
        if (nrm == 0.0 || (sig - x0 * x0) == 0.0) {
            /* already in tridiagonal form for this column */
            e[j] = x0;
            for (int i = 0; i < m; ++i) vs[IDX(i, j, n)] = 0.0;
            continue;
        }
        double alpha = (x0 >= 0.0) ? -nrm : nrm;
        v[0] = x0 - alpha;
        double vtv = sig - x0 * x0 + v[0] * v[0];
        double t = 2.0 / vtv;
        tau[j] = t;
        for (int i = 0; i < m; ++i) vs[IDX(i, j, n)] = v[i];
        /* p = t * A22 v */
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int c = 0; c < m; ++c) s += a[IDX(j + 1 + i, j + 1 + c, n)] * v[c];
            p[i] = t * s;
        }
        double K = 0.0;
        for (int i = 0; i < m; ++i) K += v[i] * p[i];
        K *= 0.5 * t;
        for (int i = 0; i < m; ++i) w[i] = p[i] - K * v[i];
        for (int c = 0; c < m; ++c)
            for (int i = 0; i < m; ++i)
                a[IDX(j + 1 + i, j + 1 + c, n)] -= v[i] * w[c] + w[i] * v[c];
        e[j] = alpha;
    }
    for (int j = 0; j < n; ++j) d[j] = a[IDX(j, j, n)];
    if (n >= 2) e[n - 2] = a[IDX(n - 1, n - 2, n)];
    if (n >= 1) e[n - 1] = 0.0;
    /* q = H_0 H_1 ... H_{n-3}, accumulated backwards */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) q[IDX(i, j, n)] = (i == j) ? 1.0 : 0.0;
    for (int j = n - 3; j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        int m = n - j - 1;
        const double *vv = vs + IDX(0, j, n);
        for (int c = j + 1; c < n; ++c) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += vv[i] * q[IDX(j + 1 + i, c, n)];
            s *= tau[j];
            for (int i = 0; i < m; ++i) q[IDX(j + 1 + i, c, n)] -= s * vv[i];
        }
    }
    free(vs);
    free(tau);
}

/*
 * Implicit QL with Wilkinson shift on the tridiagonal (d, e), accumulating
 * the rotations into z (n x n, column-major, columns are eigenvectors).
 * Bowdler, Martin, Reinsch & Wilkinson (1968); Golub & Van Loan Sec. 8.3.5.
 * Returns 0, or l+1 when eigenvalue l did not converge in 30n iterations.
 */
static int orc_tql(int n, double *d, double *e, double *z, int max_iter)
{
    if (n <= 1) return 0;
    e[n - 1] = 0.0;
    /* a subdiagonal entry is negligible relative to its neighbours or to the
       norm of T (the EISPACK tql2 test); the absolute part deflates the
       noise-level blocks a rank-deficient A leaves */
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i) {
        double t = fabs(d[i]) + fabs(e[i]);
        if (t > tnorm) tnorm = t;
    }
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= DBL_EPSILON * dd || fabs(e[m]) <= DBL_EPSILON * tnorm) break;
            }
            if (m == l) break;
            if (++iter > max_iter) return l + 1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            int underflow = 0;
            for (i = m - 1; i >= l; --i) {
                double f = s * e[i];
                double b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    underflow = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                double *zi = z + IDX(0, i, n), *zi1 = z + IDX(0, i + 1, n);
                for (int q = 0; q < n; ++q) {
                    double t = zi1[q];
                    zi1[q] = s * zi[q] + c * t;
                    zi[q] = c * zi[q] - s * t;
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

static void orc_sort_eig(int n, double *w, double *v)
{
    double *tmp = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j) if (w[j] < w[k]) k = j;
        if (k != i) {
            double t = w[i]; w[i] = w[k]; w[k] = t;
            memcpy(tmp, v + IDX(0, i, n), sizeof(double) * n);
            memcpy(v + IDX(0, i, n), v + IDX(0, k, n), sizeof(double) * n);
            memcpy(v + IDX(0, k, n), tmp, sizeof(double) * n);
        }
    }
    free(tmp);
}

/* a (n x n, lower triangle read) -> eigenvalues w ascending, vectors v. */
int orc_syev_ql(int n, const double *a, double *w, double *v)
{
    double *aa = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *e = (double *)malloc(sizeof(double) * (n + 1));
    double *work = (double *)malloc(sizeof(double) * 3 * (n + 1));
    memcpy(aa, a, sizeof(double) * (size_t)n * n);
    orc_tridiag(n, aa, w, e, v, work);
    int info = orc_tql(n, w, e, v, 30 * n);
    orc_sort_eig(n, w, v);
    free(aa); free(e); free(work);
    return info;
}

/* Cyclic Jacobi, Golub & Van Loan Alg. 8.5.3 (lower triangle read). */
int orc_syev_jacobi(int n, const double *a, double *w, double *v)
{
    double *aa = (double *)malloc(sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            aa[IDX(i, j, n)] = (i >= j) ? a[IDX(i, j, n)] : a[IDX(j, i, n)];
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) v[IDX(i, j, n)] = (i == j) ? 1.0 : 0.0;
    int info = 1;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double x = aa[IDX(i, j, n)] * aa[IDX(i, j, n)];
                tot += x;
                if (i != j) off += x;
            }
        (void)off;
        double floor_abs = 0.1 * DBL_EPSILON * sqrt(tot);
        int rotations = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = aa[IDX(p, q, n)];
                double app = aa[IDX(p, p, n)], aqq = aa[IDX(q, q, n)];
                if (fabs(apq) <= floor_abs ||
                    fabs(apq) <= 0.5 * DBL_EPSILON * sqrt(fabs(app) * fabs(aqq))) {
                    aa[IDX(p, q, n)] = 0.0;
                    aa[IDX(q, p, n)] = 0.0;
                    continue;
                }
                ++rotations;
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {   /* A <- A J */
                    double akp = aa[IDX(k, p, n)], akq = aa[IDX(k, q, n)];
                    aa[IDX(k, p, n)] = c * akp - s * akq;
                    aa[IDX(k, q, n)] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {   /* A <- J^T A */
                    double apk = aa[IDX(p, k, n)], aqk = aa[IDX(q, k, n)];
                    aa[IDX(p, k, n)] = c * apk - s * aqk;
                    aa[IDX(q, k, n)] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {   /* V <- V J */
                    double vkp = v[IDX(k, p, n)], vkq = v[IDX(k, q, n)];
                    v[IDX(k, p, n)] = c * vkp - s * vkq;
                    v[IDX(k, q, n)] = s * vkp + c * vkq;
                }
            }
        if (rotations == 0) { info = 0; break; }
    }
    for (int i = 0; i < n; ++i) w[i] = aa[IDX(i, i, n)];
    orc_sort_eig(n, w, v);
    free(aa);
    return info;
}

/* ------------------------------------------------------------------------ */
/* letkf_core                                                                */
/* ------------------------------------------------------------------------ */

/*
 * Eigenvalues of A (ascending) -> eigenvalues of A + (k-1)/rho I.
 * letkf_core.F90:107-113: eival = eival*1 + (k-1)/inflation_factor.
 * ORC_GUARD first lifts a negative spectrum: if the smallest eigenvalue of
 * A is below 0 every eigenvalue is shifted up by that deficit -- the guard of
 * the reference's reduced-precision cores (Sunway letkf_core_half_master.c:
 * 157-162; letkf_core_mpal.F90:113-118 is the same test written after the
 * (k-1)/rho shift, identical for rho = 1).
 */
static void orc_shift_eigenvalues(int k, double *eival, double inflation_factor, int flags)
{
    if ((flags & ORC_GUARD) && eival[0] < 0.0) {
        double mn = eival[0];
        for (int i = 0; i < k; ++i) eival[i] = eival[i] - mn;
    }
    for (int i = 0; i < k; ++i)
        eival[i] = eival[i] * 1.0 + (double)(k - 1) / inflation_factor;
}

/*
 * letkf_core.F90:14-233.  hxens_in is (nanals, nobsl) column-major; trans is
 * (nanals, nanals) column-major.  flags: ORC_EIG_JACOBI selects the Jacobi
 * eigensolver; ORC_GUARD applies the letkf_core_mpal eigenvalue guard
 * (letkf_core_mpal.F90:73-76,113-118): (k-1)/rho is added to the diagonal
 * before the eigensolve and, if the smallest eigenvalue is below k-1, all
 * eigenvalues are shifted up by the deficit.  Returns the eigensolver status
 * (0 = ok), which the reference only prints (letkf_core.F90:119-122).
 */
int orc_letkf_core(int nanals, int nobsl, const double *hxens_in, const float *rdiaginv,
                   const double *dep, const double *rloc, double *trans, double *overline_w,
                   double inflation_factor, int flags)
{
    const int k = nanals;
    const size_t kk = (size_t)k * k;
    double *hx = (double *)malloc(sizeof(double) * ((size_t)k * (nobsl > 0 ? nobsl : 1)));
    double *rr = (double *)malloc(sizeof(double) * (nobsl > 0 ? nobsl : 1));
    double *work1 = (double *)calloc(kk, sizeof(double));
    double *eivec = (double *)malloc(sizeof(double) * kk);
    double *eival = (double *)malloc(sizeof(double) * k);
    double *pa = (double *)malloc(sizeof(double) * kk);
    double *tmp = (double *)malloc(sizeof(double) * kk);

    /* :58-69 row scaling by sqrt(rdiaginv * rloc) */
    for (int j = 0; j < nobsl; ++j) {
        double t = (double)rdiaginv[j] * rloc[j];
        t = sqrt(t);
        rr[j] = t;
        for (int i = 0; i < k; ++i) hx[IDX(i, j, k)] = hxens_in[IDX(i, j, k)] * t;
    }
    /* :74 dgemm('n','t') -> work1 = hx hx^T (full matrix, sum over obs) */
    for (int j = 0; j < nobsl; ++j) {
        const double *y = hx + IDX(0, j, k);
        for (int c = 0; c < k; ++c) {
            double yc = y[c];
            double *col = work1 + IDX(0, c, k);
            for (int r = 0; r < k; ++r) col[r] += y[r] * yc;
        }
    }
    /* :98 dsyevd('V','L') */
    int info = (flags & ORC_EIG_JACOBI) ? orc_syev_jacobi(k, work1, eival, eivec)
                                        : orc_syev_ql(k, work1, eival, eivec);
    orc_shift_eigenvalues(k, eival, inflation_factor, flags);
    /* :129-143 pa = (eivec / eival) eivec^T */
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) tmp[IDX(i, j, k)] = eivec[IDX(i, j, k)] / eival[j];
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += tmp[IDX(i, l, k)] * eivec[IDX(j, l, k)];
            pa[IDX(i, j, k)] = s;
        }
    /* :146-148 hx *= rr  (now hxens_in * rinv * rloc) */
    for (int j = 0; j < nobsl; ++j)
        for (int i = 0; i < k; ++i) hx[IDX(i, j, k)] *= rr[j];
    /* :164 work2 = pa hx ; :190-195 overline_w = work2 dep */
    for (int i = 0; i < k; ++i) overline_w[i] = 0.0;
    {
        double *w2 = (double *)malloc(sizeof(double) * k);
        for (int j = 0; j < nobsl; ++j) {
            for (int i = 0; i < k; ++i) {
                double s = 0.0;
                for (int l = 0; l < k; ++l) s += pa[IDX(i, l, k)] * hx[IDX(l, j, k)];
                w2[i] = s;
            }
            for (int i = 0; i < k; ++i) overline_w[i] += w2[i] * dep[j];
        }
        free(w2);
    }
    /* :197-210 trans = (eivec * sqrt((k-1)/eival)) eivec^T */
    for (int j = 0; j < k; ++j) {
        double rho = sqrt((double)(k - 1) / eival[j]);
        for (int i = 0; i < k; ++i) tmp[IDX(i, j, k)] = eivec[IDX(i, j, k)] * rho;
    }
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += tmp[IDX(i, l, k)] * eivec[IDX(j, l, k)];
            trans[IDX(i, j, k)] = s;
        }
    /* :221-225 trans(i,j) += overline_w(i) */
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) trans[IDX(i, j, k)] += overline_w[i];

    free(hx); free(rr); free(work1); free(eivec); free(eival); free(pa); free(tmp);
    return info;
}

/*
 * The transform half of letkf_core (letkf_core.F90:97-225) from the
 * obs-space products: a = Yb^T R^-1 rho Yb (k x k, column-major, lower
 * triangle read) and b = Yb^T R^-1 rho d.  overline_w = Pa b, which is
 * letkf_core.F90:164,190-195 with the sums regrouped (the Sunway kernel
 * computes it this way, letkf_core_half_slave.c:52).  flags as orc_letkf_core.
 */
int orc_letkf_transform(int k, const double *a, const double *b, double inflation_factor,
                        int flags, double *trans, double *overline_w)
{
    const size_t kk = (size_t)k * k;
    double *work1 = (double *)malloc(sizeof(double) * kk);
    double *eivec = (double *)malloc(sizeof(double) * kk);
    double *eival = (double *)malloc(sizeof(double) * k);
    double *tmp = (double *)malloc(sizeof(double) * kk);
    double *vb = (double *)malloc(sizeof(double) * k);
    memcpy(work1, a, sizeof(double) * kk);
    int info = (flags & ORC_EIG_JACOBI) ? orc_syev_jacobi(k, work1, eival, eivec)
                                        : orc_syev_ql(k, work1, eival, eivec);
    orc_shift_eigenvalues(k, eival, inflation_factor, flags);
    /* overline_w = V diag(1/eival) V^T b */
    for (int j = 0; j < k; ++j) {
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += eivec[IDX(i, j, k)] * b[i];
        vb[j] = s / eival[j];
    }
    for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += eivec[IDX(i, j, k)] * vb[j];
        overline_w[i] = s;
    }
    for (int j = 0; j < k; ++j) {
        double rho = sqrt((double)(k - 1) / eival[j]);
        for (int i = 0; i < k; ++i) tmp[IDX(i, j, k)] = eivec[IDX(i, j, k)] * rho;
    }
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += tmp[IDX(i, l, k)] * eivec[IDX(j, l, k)];
            trans[IDX(i, j, k)] = s + overline_w[i];
        }
    free(work1); free(eivec); free(eival); free(tmp); free(vb);
    return info;
}

/* ------------------------------------------------------------------------ */
/* Ensemble update                                                           */
/* ------------------------------------------------------------------------ */

/*
 * xb_2_xa, letkf_compute.F90:259-329, for one block of m contiguous state
 * elements: xa(:,c) = trans^T xb(:,c) + axb(c); axa(c) = mean_e xa(e,c);
 * xa(:,c) -= axa(c).
 */
void orc_xb_2_xa(int k, int m, const double *xb, const double *axb, double *xa, double *axa,
                 const double *trans)
{
    for (int c = 0; c < m; ++c) {
        const double *xbc = xb + (size_t)c * k;
        double *xac = xa + (size_t)c * k;
        for (int e = 0; e < k; ++e) {
            double s = 0.0;
            for (int f = 0; f < k; ++f) s += trans[IDX(f, e, k)] * xbc[f];
            xac[e] = s;
        }
        for (int e = 0; e < k; ++e) xac[e] = xac[e] + axb[c];
        double a = 0.0;
        for (int e = 0; e < k; ++e) a += xac[e];
        a = a / k;
        axa[c] = a;
        for (int e = 0; e < k; ++e) xac[e] -= a;
    }
}

/* ------------------------------------------------------------------------ */
/* Loops over local analyses                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    const orc_problem *pb;
    const double *xb, *axb;
    double *xa, *axa;
    double *trans_out, *w_out;   /* optional per-analysis outputs */
    int *nobsl_out;
    int *status_out;
    int a_begin, a_end;          /* analysis index range [begin, end) */
    const int *a_list;           /* optional explicit analysis list */
    double *cube_trans;          /* overlapped cubes: every cube's trans */
    int *cube_status;
    int rc;
} orc_job;

/* Cube start stride: cube_n - overlap (letkf_compute.F90:197,205), or
   cube_n without overlap (:53,61). */
static int orc_cube_stride(const orc_problem *pb)
{
    int cn = pb->cube_n > 0 ? pb->cube_n : 1;
    return pb->overlap > 0 ? cn - pb->overlap : cn;
}

/* cubes along n grid points: starts 1, 1 + s, ... <= n */
static int orc_cubes_along(int n, int s) { return (n - 1) / s + 1; }

static int orc_n_analyses_h(const orc_problem *pb)
{
    int s = orc_cube_stride(pb);
    return orc_cubes_along(pb->nlat, s) * orc_cubes_along(pb->nlon, s);
}

int orc_n_analyses(const orc_problem *pb)
{
    int na = orc_n_analyses_h(pb);
    if (pb->vert_dist > 0.0) {
        int cz = pb->cube_nz > 0 ? pb->cube_nz : pb->nz;
        na *= (pb->nz + cz - 1) / cz;
    }
    return na;
}

/*
 * Analysis a = (lon cube index jc) * nlat_cubes + (lat cube index ic), the
 * column-major order of xb(ens, n, lat, lon).  For the vertical loop each
 * horizontal analysis owns nz_chunks consecutive indices.
 */
static size_t col_off(const orc_problem *pb, int ilat, int ilon)
{
    /* xb(ens, n, 1-h:nlat+h, 1-h:nlon+h); ilat, ilon are 0-based interior */
    int h = pb->halo;
    size_t lat_ext = (size_t)pb->nlat + 2 * h;
    return (((size_t)(ilon + h) * lat_ext) + (size_t)(ilat + h)) * (size_t)pb->n;
}

static int orc_run_horizontal(orc_job *jb, int ah, double *yb, double *dis, float *rdi,
                              double *rloc, double *vco, double *trans, double *w, int cap)
{
    const orc_problem *pb = jb->pb;
    const int k = pb->k, cn = pb->cube_n > 0 ? pb->cube_n : 1;
    const int cs = orc_cube_stride(pb);
    const int nlat_c = orc_cubes_along(pb->nlat, cs);
    int ic = ah % nlat_c, jc = ah / nlat_c;
    /* letkf_compute.F90:53-65 and :195-209 (1-based, clipped), centre
       (sj+ej)/2, (si+ei)/2 */
    int sj = ic * cs + 1, ej = sj + cn - 1; if (ej > pb->nlat) ej = pb->nlat;
    int si = jc * cs + 1, ei = si + cn - 1; if (ei > pb->nlon) ei = pb->nlon;
    double clat = (double)pb->grid_lat[(sj + ej) / 2 - 1];
    double clon = (double)pb->grid_lon[(si + ei) / 2 - 1];
    int nobsl = orc_gen_yb_dis(pb->obs, k, clat, clon, pb->local_dist, pb->gc_filter,
                               pb->obs_polar, pb->max_obs, cap, yb, dis, rdi, rloc, vco);
    if (nobsl < 0) return -2;
    if (pb->vert_dist <= 0.0) {
        int st = orc_letkf_core(k, nobsl, yb, rdi, dis, rloc, trans, w, pb->inflation_factor,
                                pb->core_flags);
        if (jb->trans_out) memcpy(jb->trans_out + (size_t)ah * k * k, trans, sizeof(double) * k * k);
        if (jb->w_out) memcpy(jb->w_out + (size_t)ah * k, w, sizeof(double) * k);
        if (jb->nobsl_out) jb->nobsl_out[ah] = nobsl;
        if (jb->status_out) jb->status_out[ah] = st;
        if (pb->overlap > 0) {
            /* letkf_compute: the cube's trans is added to its columns by
               xb_add2_xa after every cube is solved (orc_overlap_update) */
            memcpy(jb->cube_trans + (size_t)ah * k * k, trans, sizeof(double) * k * k);
            jb->cube_status[ah] = st;
            return 0;
        }
        /* xb_2_xa over the cube: columns are not contiguous across lon, so
           update one (lat-run, lon) strip at a time */
        for (int lo = si; lo <= ei; ++lo) {
            size_t off = col_off(pb, sj - 1, lo - 1);
            int mcols = (ej - sj + 1) * pb->n;
            orc_xb_2_xa(k, mcols, jb->xb + off * k, jb->axb + off, jb->xa + off * k,
                        jb->axa + off, trans);
        }
        return 0;
    }
    /* letkf_compute_nooverlap_vertical, letkf_compute.F90:436-537 */
    const int nz = pb->nz, cz = pb->cube_nz > 0 ? pb->cube_nz : nz;
    const int nzc = (nz + cz - 1) / cz;
    double *rl = (double *)malloc(sizeof(double) * (nobsl > 0 ? nobsl : 1));
    int cj = (sj + ej) / 2 - 1, ci = (si + ei) / 2 - 1;
    for (int kc = 0; kc < nzc; ++kc) {
        int sk = kc * cz + 1, ek = sk + cz - 1; if (ek > nz) ek = nz;
        int center_k = (sk + ek) / 2;
        double lnp_c;
        if (pb->vcoord_mode == ORC_VCOORD_LEVEL) {
            lnp_c = (double)center_k;
        } else {
            /* ps2alnp, obs_make.F90:418-437 with the hybrid p(ps) of
               vert_coord_mod.F90:556-563 */
            double ps = jb->axb[col_off(pb, cj, ci) + pb->ps_index];
            double ph0 = pb->hyai[center_k - 1] * (pb->p0 - pb->ptop) +
                         pb->hybi[center_k - 1] * (ps - pb->ptop) + pb->ptop;
            double ph1 = pb->hyai[center_k] * (pb->p0 - pb->ptop) +
                         pb->hybi[center_k] * (ps - pb->ptop) + pb->ptop;
            lnp_c = log(0.5 * (ph0 + ph1));
        }
        orc_vertical_localization(rloc, rl, nobsl, lnp_c, vco, pb->vert_dist,
                                  pb->gc_filter_vertical);
        int st = orc_letkf_core(k, nobsl, yb, rdi, dis, rl, trans, w, pb->inflation_factor,
                                pb->core_flags);
        int a = ah * nzc + kc;
        if (jb->trans_out) memcpy(jb->trans_out + (size_t)a * k * k, trans, sizeof(double) * k * k);
        if (jb->w_out) memcpy(jb->w_out + (size_t)a * k, w, sizeof(double) * k);
        if (jb->nobsl_out) jb->nobsl_out[a] = nobsl;
        if (jb->status_out) jb->status_out[a] = st;
        /* :501-530 per-variable update of this chunk's levels */
        for (int v = 0; v < pb->n_var; ++v) {
            int start = pb->var_start[v], lev = pb->var_lev[v];
            int l0, l1;
            if (lev == 1) {
                if (ek != nz) continue;
                l0 = start; l1 = start;
            } else {
                l0 = start + sk - 1; l1 = start + ek - 1;
            }
            for (int lo = si; lo <= ei; ++lo)
                for (int la = sj; la <= ej; ++la) {
                    size_t off = col_off(pb, la - 1, lo - 1);
                    for (int nk = l0; nk <= l1; ++nk) {
                        const double *xbc = jb->xb + (off + nk - 1) * k;
                        double *xac = jb->xa + (off + nk - 1) * k;
                        double ab = jb->axb[off + nk - 1];
                        for (int e = 0; e < k; ++e) {
                            double s = 0.0;
                            for (int f = 0; f < k; ++f) s += xbc[f] * trans[IDX(f, e, k)];
                            xac[e] = ab + s;
                        }
                    }
                }
        }
    }
    free(rl);
    /* :539-552 re-centre (done per column here; columns are independent) */
    for (int lo = si; lo <= ei; ++lo)
        for (int la = sj; la <= ej; ++la) {
            size_t off = col_off(pb, la - 1, lo - 1);
            for (int v = 0; v < pb->n; ++v) {
                double *xac = jb->xa + (off + v) * k;
                double a = 0.0;
                for (int e = 0; e < k; ++e) a += xac[e];
                a = a / k;
                jb->axa[off + v] = a;
                for (int e = 0; e < k; ++e) xac[e] -= a;
            }
        }
    return 0;
}

static void *orc_worker(void *arg)
{
    orc_job *jb = (orc_job *)arg;
    const orc_problem *pb = jb->pb;
    const int k = pb->k;
    int cap = pb->obs->n_rows > 0 ? pb->obs->n_rows : 1;
    if (pb->max_obs > 0 && pb->max_obs < cap) cap = pb->max_obs;
    double *yb = (double *)malloc(sizeof(double) * (size_t)k * cap);
    double *dis = (double *)malloc(sizeof(double) * cap);
    float *rdi = (float *)malloc(sizeof(float) * cap);
    double *rloc = (double *)malloc(sizeof(double) * cap);
    double *vco = (double *)malloc(sizeof(double) * cap);
    double *trans = (double *)malloc(sizeof(double) * k * k);
    double *w = (double *)malloc(sizeof(double) * k);
    jb->rc = 0;
    for (int t = jb->a_begin; t < jb->a_end; ++t) {
        int ah = jb->a_list ? jb->a_list[t] : t;
        int rc = orc_run_horizontal(jb, ah, yb, dis, rdi, rloc, vco, trans, w, cap);
        if (rc) { jb->rc = rc; break; }
    }
    free(yb); free(dis); free(rdi); free(rloc); free(vco); free(trans); free(w);
    return NULL;
}

/*
 * letkf_compute's accumulation (letkf_compute.F90:211-255): xa = 0 and axa = 0
 * everywhere (halo included, :191-193), then for each cube in loop order
 * (lon cubes outer, lat cubes inner) xb_add2_xa (:332-382) adds to each of
 * its columns, per state element v,
 *     dist(e) = axb(v) + dot_product(xb(:, v), trans(:, e)),
 *     mean = sum_e dist(e) / ens,  axa(v) += mean,  xa(e, v) += dist(e) - mean,
 * and cube_count of the column grows by one; finally interior columns are
 * divided by cube_count (:248-255).  A column receives its cubes' terms in loop
 * order whether the cubes are visited once (the reference) or gathered per
 * column (here, lon strips over threads), so the sums are the same.
 */
typedef struct {
    const orc_problem *pb;
    const double *xb, *axb, *trans;
    double *xa, *axa;
    int lon_begin, lon_end;
} orc_ovl_job;

static void *orc_overlap_worker(void *arg)
{
    orc_ovl_job *jb = (orc_ovl_job *)arg;
    const orc_problem *pb = jb->pb;
    const int k = pb->k, n = pb->n, cn = pb->cube_n, cs = orc_cube_stride(pb);
    const int nlat_c = orc_cubes_along(pb->nlat, cs);
    const int nlon_c = orc_cubes_along(pb->nlon, cs);
    double *dist = (double *)malloc(sizeof(double) * k);
    for (int lo = jb->lon_begin; lo < jb->lon_end; ++lo) {
        /* lon cubes jc with jc*cs <= lo <= jc*cs + cn - 1 (0-based) */
        int jc0 = lo - cn + 1 <= 0 ? 0 : (lo - cn + 1 + cs - 1) / cs, jc1 = lo / cs;
        if (jc1 > nlon_c - 1) jc1 = nlon_c - 1;
        for (int la = 0; la < pb->nlat; ++la) {
            int ic0 = la - cn + 1 <= 0 ? 0 : (la - cn + 1 + cs - 1) / cs, ic1 = la / cs;
            if (ic1 > nlat_c - 1) ic1 = nlat_c - 1;
            size_t off = col_off(pb, la, lo);
            double count = 0.0;
            for (int jc = jc0; jc <= jc1; ++jc)
                for (int ic = ic0; ic <= ic1; ++ic) {
                    const double *tr = jb->trans + (size_t)(jc * nlat_c + ic) * k * k;
                    for (int v = 0; v < n; ++v) {
                        const double *xbc = jb->xb + (off + v) * k;
                        double *xac = jb->xa + (off + v) * k;
                        double ab = jb->axb[off + v];
                        double mean = 0.0;
                        for (int e = 0; e < k; ++e) {
                            double s = 0.0;
                            for (int f = 0; f < k; ++f) s += xbc[f] * tr[IDX(f, e, k)];
                            dist[e] = ab + s;
                            mean += dist[e];
                        }
                        mean = mean / k;
                        jb->axa[off + v] += mean;
                        for (int e = 0; e < k; ++e) xac[e] += dist[e] - mean;
                    }
                    count += 1.0;
                }
            for (int v = 0; v < n; ++v) {
                double *xac = jb->xa + (off + v) * k;
                for (int e = 0; e < k; ++e) xac[e] = xac[e] / count;
                jb->axa[off + v] = jb->axa[off + v] / count;
            }
        }
    }
    free(dist);
    return NULL;
}

static void orc_overlap_update(const orc_problem *pb, const double *xb, const double *axb,
                               const double *trans, double *xa, double *axa, int nthreads)
{
    const size_t cols = (size_t)(pb->nlat + 2 * pb->halo) * (pb->nlon + 2 * pb->halo);
    memset(xa, 0, sizeof(double) * cols * pb->n * pb->k);
    memset(axa, 0, sizeof(double) * cols * pb->n);
    if (nthreads > pb->nlon) nthreads = pb->nlon;
    if (nthreads < 1) nthreads = 1;
    orc_ovl_job *jobs = (orc_ovl_job *)calloc(nthreads, sizeof(orc_ovl_job));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        orc_ovl_job *jb = &jobs[t];
        jb->pb = pb; jb->xb = xb; jb->axb = axb; jb->trans = trans; jb->xa = xa; jb->axa = axa;
        jb->lon_begin = (int)((long long)pb->nlon * t / nthreads);
        jb->lon_end = (int)((long long)pb->nlon * (t + 1) / nthreads);
    }
    if (nthreads == 1) {
        orc_overlap_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, orc_overlap_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    free(jobs); free(th);
}

/*
 * Runs letkf_compute_nooverlap (vert_dist <= 0) or
 * letkf_compute_nooverlap_vertical (vert_dist > 0) over horizontal analyses
 * [h_begin, h_end) (or the explicit list h_list of n_list entries when
 * non-NULL) with nthreads POSIX threads.  Optional outputs: trans (k*k per
 * analysis), overline_w (k per analysis), nobsl, status.
 */
int orc_letkf_compute(const orc_problem *pb, const double *xb, const double *axb, double *xa,
                      double *axa, double *trans_out, double *w_out, int *nobsl_out,
                      int *status_out, const int *h_list, int n_list, int nthreads)
{
    const int ovl = pb->overlap > 0;
    if (ovl && (h_list || pb->vert_dist > 0.0 || pb->overlap >= pb->cube_n)) return -3;
    int nh = h_list ? n_list : orc_n_analyses_h(pb);
    double *cube_trans = ovl ? (double *)malloc(sizeof(double) * (size_t)nh * pb->k * pb->k) : NULL;
    int *cube_status = ovl ? (int *)calloc(nh > 0 ? nh : 1, sizeof(int)) : NULL;
    const int nthreads_all = nthreads < 1 ? 1 : nthreads;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nh) nthreads = nh > 0 ? nh : 1;
    orc_job *jobs = (orc_job *)calloc(nthreads, sizeof(orc_job));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        orc_job *jb = &jobs[t];
        jb->pb = pb; jb->xb = xb; jb->axb = axb; jb->xa = xa; jb->axa = axa;
        jb->trans_out = trans_out; jb->w_out = w_out; jb->nobsl_out = nobsl_out;
        jb->status_out = status_out;
        jb->a_list = h_list;
        jb->cube_trans = cube_trans;
        jb->cube_status = cube_status;
        jb->a_begin = (int)((long long)nh * t / nthreads);
        jb->a_end = (int)((long long)nh * (t + 1) / nthreads);
    }
    if (nthreads == 1) {
        orc_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, orc_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    int rc = 0;
    for (int t = 0; t < nthreads; ++t) if (jobs[t].rc) rc = jobs[t].rc;
    free(jobs); free(th);
    if (ovl && rc == 0) orc_overlap_update(pb, xb, axb, cube_trans, xa, axa, nthreads_all);
    free(cube_trans); free(cube_status);
    return rc;
}

/*
 * run_da_parallel epilogue, module_da_parallel.F90:219-261: for interior
 * columns with |grid_lat| <= polar_lat the member is
 * gamma1*(axb+xb) + gamma2*(axa+xa), otherwise xb+axb; spreads are RMS over
 * members with divisor ens.  Halo columns of xa_out are set to 0.
 */
void orc_epilogue(const orc_problem *pb, const double *xb, const double *axb, const double *xa,
                  const double *axa, double polar_lat, double gamma1, double gamma2,
                  double *xa_out, double *spread_b, double *spread_a)
{
    const int k = pb->k, n = pb->n, h = pb->halo;
    const int lat_ext = pb->nlat + 2 * h, lon_ext = pb->nlon + 2 * h;
    for (int lo = 0; lo < lon_ext; ++lo)
        for (int la = 0; la < lat_ext; ++la) {
            size_t off = ((size_t)lo * lat_ext + la) * n;
            int interior = lo >= h && lo < pb->nlon + h && la >= h && la < pb->nlat + h;
            for (int v = 0; v < n; ++v) {
                const double *b = xb + (off + v) * k;
                const double *a = xa + (off + v) * k;
                double *o = xa_out + (off + v) * k;
                if (!interior) {
                    for (int e = 0; e < k; ++e) o[e] = 0.0;
                    continue;
                }
                size_t so = ((size_t)(lo - h) * pb->nlat + (la - h)) * n + v;
                double ab = axb[off + v], aa = axa[off + v];
                double sb = 0.0, sa = 0.0;
                if (fabs((double)pb->grid_lat[la - h]) <= polar_lat) {
                    for (int e = 0; e < k; ++e)
                        o[e] = gamma1 * (ab + b[e]) + gamma2 * (aa + a[e]);
                    for (int e = 0; e < k; ++e) {
                        sb += b[e] * b[e];
                        sa += (o[e] - aa) * (o[e] - aa);
                    }
                } else {
                    for (int e = 0; e < k; ++e) o[e] = b[e] + ab;
                    for (int e = 0; e < k; ++e) { sb += b[e] * b[e]; sa += b[e] * b[e]; }
                }
                spread_b[so] = sqrt(sb / (double)(float)k);
                spread_a[so] = sqrt(sa / (double)(float)k);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* observation operator (gen_obs_struct, letkf_obs.F90:16-187)               */
/* ------------------------------------------------------------------------ */

/* find_obs_grid (letkf_obs.F90:192-248): first 1-based index with
 * grid >= value over the closed grid (n + 1 entries: the last is 90 / 360);
 * the weight of the point before it, with the spacing taken in real(4)
 * (grid_lat(tx) - grid_lat(tx-1) is a real(4) expression, :234). */
static void orc_find_obs_grid(const float *g, int n, float g_close, double v, int *t, double *w)
{
    int tx = 0;
    for (int i = 1; i <= n + 1; ++i) {
        double gi = (i <= n) ? (double)g[i - 1] : (double)g_close;
        if (gi >= v) { tx = i; break; }
    }
    *t = tx;
    if (tx > 1) {
        float ga = g[tx - 2];
        float gb = (tx <= n) ? g[tx - 1] : g_close;
        double d = (double)(float)(gb - ga);
        *w = (d - v + (double)ga) / d;
    } else {
        *w = 0.0;
    }
}

/* da_hybrid_coord_calc_ph_lev (vert_coord_mod.F90:556-561), 1-based level */
static double orc_ph_lev(const orc_problem *pb, int lev, double phs)
{
    return pb->hyai[lev - 1] * (pb->p0 - pb->ptop) + pb->hybi[lev - 1] * (phs - pb->ptop) + pb->ptop;
}

int orc_obs_operator(const orc_problem *pb, const double *xb, const double *axb, int n_sites,
                     const double *site_lat, const double *site_lon, const int *site_row_ptr,
                     const int *row_var, const double *row_ph, const double *yo, double *yb,
                     double *dep)
{
    const int k = pb->k, n = pb->n, nz = pb->nz, h = pb->halo;
    if (h < 1 || !pb->hyai || !pb->hybi || pb->ps_index < 0 || pb->ps_index >= n) return -1;
    const size_t lat_ext = (size_t)pb->nlat + 2 * h;
    const size_t np = (size_t)nz + 1;
    double *ph = (double *)malloc(sizeof(double) * np * (size_t)k);
    double *lev = (double *)malloc(sizeof(double) * np);
    int *jpos = (int *)malloc(sizeof(int) * (size_t)k);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)k);
    if (!ph || !lev || !jpos || !tmp) { free(ph); free(lev); free(jpos); free(tmp); return -1; }
    for (int s = 0; s < n_sites; ++s) {
        int tx, ty;
        double ax, ay;
        orc_find_obs_grid(pb->grid_lat, pb->nlat, 90.0f, site_lat[s], &tx, &ax);
        orc_find_obs_grid(pb->grid_lon, pb->nlon, 360.0f, site_lon[s], &ty, &ay);
        /* the 2 x 2 points (tx-1 | tx) x (ty-1 | ty), 1-based with the halo */
        size_t pt[2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                pt[a][b] = ((size_t)(ty - 2 + b + h) * lat_ext + (size_t)(tx - 2 + a + h)) * n;
        /* bilinear value of state element el for member e (obs_make.F90:284-293) */
#define ORC_BIL(el, e) ( ((xb[(pt[0][0] + (el)) * k + (e)] + axb[pt[0][0] + (el)]) * ax +          \
                          (xb[(pt[1][0] + (el)) * k + (e)] + axb[pt[1][0] + (el)]) * (1 - ax)) * ay + \
                         ((xb[(pt[0][1] + (el)) * k + (e)] + axb[pt[0][1] + (el)]) * ax +          \
                          (xb[(pt[1][1] + (el)) * k + (e)] + axb[pt[1][1] + (el)]) * (1 - ax)) * (1 - ay) )
        /* member pressure profiles (ps2ph_3d_interp, obs_make.F90:345-381) */
        for (int e = 0; e < k; ++e) {
            const double ps = ORC_BIL(pb->ps_index, e);
            for (int l = 1; l <= nz + 1; ++l) lev[l - 1] = orc_ph_lev(pb, l, ps);
            double *p = ph + (size_t)e * np;
            for (int l = 0; l < nz; ++l) p[l] = 0.5 * (lev[l] + lev[l + 1]);
            p[nz] = p[nz - 1];
            jpos[e] = 1;
        }
        const int r0 = site_row_ptr[s], r1 = site_row_ptr[s + 1];
        for (int r = r0; r < r1; ++r) {
            const int v = row_var[r];
            const int vs = pb->var_start[v] - 1;
            for (int e = 0; e < k; ++e) {
                const double *p = ph + (size_t)e * np;
                /* ph_pos search, continuing from the previous level (:251-258) */
                int j = jpos[e];
                while (j <= nz && row_ph[r] > p[j - 1]) ++j;
                jpos[e] = j;
                const int pos = j - 1;
                if (pb->var_lev[v] > 1) {
                    if (pos >= 1 && pos < nz) {
                        const int tl = vs + pos - 1;
                        const double t3 = (log(p[pos]) - log(row_ph[r])) / (log(p[pos]) - log(p[pos - 1]));
                        const double rA = ORC_BIL(tl, e), rB = ORC_BIL(tl + 1, e);
                        tmp[e] = rA * t3 + rB * (1 - t3);
                    } else {
                        const int tl = (pos == 0) ? vs : vs + nz - 1;
                        tmp[e] = ORC_BIL(tl, e);
                    }
                } else {
                    tmp[e] = ORC_BIL(vs, e);
                }
            }
            /* centring (letkf_obs.F90:158-170) */
            double sum = 0.0;
            for (int e = 0; e < k; ++e) sum += tmp[e];
            sum = sum / k;
            dep[r] = yo[r] - sum;
            for (int e = 0; e < k; ++e) yb[(size_t)r * k + e] = tmp[e] - sum;
        }
#undef ORC_BIL
    }
    free(ph); free(lev); free(jpos); free(tmp);
    return 0;
}
