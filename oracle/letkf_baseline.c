/*
 * letkf_baseline.c -- the reference's x86 LETKF analysis step on host cores,
 * for the CPU baseline that bench.py times beside the GPU engine.
 * TEST/BENCH INFRASTRUCTURE ONLY: never part of the product.
 *
 * What runs, per the reference:
 *  - the ranks: P threads, each owning one block of the two-way split
 *    MPI_Dims_create(P, 2) gives (module_da_parallel.F90:67-105), blocks
 *    sized as pro_info_get_indx (pro_info.F90:96-116);
 *  - each rank's candidate obs: the sites in the blocks comm_info_init picks
 *    (letkf_init.F90:100-160): north/south while the meridian distance from
 *    the block's edge row to the other block's near edge is <= local_dist,
 *    west/east while nearest_dist (letkf_init.F90:288-341) is <= local_dist;
 *  - per column of the block: gen_yb_dis's brute-force scan of the candidates
 *    (latlon_dist_r8, schr2; letkf_yb.F90:3-150), letkf_core
 *    (letkf_core.F90:58-225: dgemm('n','t') for the products, dsyevd, dgemm for
 *    Pa, Pa Yb R^-1 and T) and xb_2_xa (letkf_compute.F90:259-329: dgemm('t','n')
 *    then the mean/perturbation split).
 * dgemm/dsyevd come from MKL (the reference's x86 provider,
 * x86/CMakeLists.txt:118-123), loaded at run time with sequential threading;
 * without MKL the oracle's own Householder + QL eigensolver and loops run.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "letkf_oracle.h"

typedef void (*dgemm_fn)(const char *, const char *, const int *, const int *, const int *,
                         const double *, const double *, const int *, const double *,
                         const int *, const double *, double *, const int *);
typedef void (*dsyevd_fn)(const char *, const char *, const int *, double *, const int *,
                          double *, double *, const int *, int *, const int *, int *);

static dgemm_fn p_dgemm = NULL;
static dsyevd_fn p_dsyevd = NULL;
static char g_provider[160] = "in-repo Householder + QL (no MKL found)";
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void load_blas(void)
{
    const char *names[] = {"libmkl_rt.so.2", "libmkl_rt.so.1", "libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so.2", "/opt/conda/lib/libmkl_rt.so.1",
                           "/opt/conda/lib/libmkl_rt.so"};
    for (size_t i = 0; i < sizeof(names) / sizeof(names[0]); ++i) {
        void *h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
        if (set_layer) set_layer(1);            /* MKL_THREADING_SEQUENTIAL */
        void (*set_nt)(int) = (void (*)(int))dlsym(h, "MKL_Set_Num_Threads");
        if (set_nt) set_nt(1);
        p_dgemm = (dgemm_fn)dlsym(h, "dgemm_");
        p_dsyevd = (dsyevd_fn)dlsym(h, "dsyevd_");
        if (p_dgemm && p_dsyevd) {
            snprintf(g_provider, sizeof(g_provider), "MKL (%s, sequential) dgemm/dsyevd", names[i]);
            return;
        }
        p_dgemm = NULL;
        p_dsyevd = NULL;
    }
}

const char *orc_baseline_provider(void)
{
    pthread_once(&g_once, load_blas);
    return g_provider;
}

/* MPI_Dims_create(p, 2): the most balanced factorisation, larger first */
static void dims_create(int p, int *d0, int *d1)
{
    int best = 1;
    for (int a = 1; a * a <= p; ++a)
        if (p % a == 0) best = a;
    *d0 = p / best;
    *d1 = best;
}

/* pro_info_get_indx: 1-based start/end of block b (1-based) of n points in nb blocks */
static void block_range(int b, int n, int nb, int *s, int *e)
{
    int sec = n / nb, res = n % nb;
    if (b - 1 < res) {
        *s = (b - 1) * (sec + 1) + 1;
        *e = *s + sec;
    } else {
        *s = (b - 1) * sec + res + 1;
        *e = *s + sec - 1;
    }
}

/* grid value at 1-based index i, closed past the end (pro_info.F90:77-83) */
static double glat_at(const orc_problem *pb, int i)
{
    return i > pb->nlat ? 90.0 : (double)pb->grid_lat[i - 1];
}
static double glon_at(const orc_problem *pb, int i)
{
    return i > pb->nlon ? 360.0 : (double)pb->grid_lon[i - 1];
}

typedef struct {
    const orc_problem *pb;
    int nbx, nby;             /* latitude blocks, longitude blocks */
} split_t;

/* nearest_dist (letkf_init.F90:288-341) between blocks (bx1, by1), (bx2, by2) */
static double nearest_dist(const split_t *sp, int bx1, int by1, int bx2, int by2)
{
    const orc_problem *pb = sp->pb;
    int xs1, xe1, ys1, ye1, xs2, xe2, ys2, ye2, tx1, tx2;
    block_range(bx1, pb->nlat, sp->nbx, &xs1, &xe1);
    block_range(by1, pb->nlon, sp->nby, &ys1, &ye1);
    block_range(bx2, pb->nlat, sp->nbx, &xs2, &xe2);
    block_range(by2, pb->nlon, sp->nby, &ys2, &ye2);
    if (bx1 > bx2) { tx1 = xs1; tx2 = xe2 + 1; }
    else if (bx1 < bx2) { tx1 = xe1 + 1; tx2 = xs2; }
    else if (bx1 > sp->nbx / 2) { tx1 = xe1 + 1; tx2 = xe2 + 1; }
    else { tx1 = xs1; tx2 = xs2; }
    double lat1 = glat_at(pb, tx1), lat2 = glat_at(pb, tx2);
    if (fabs(lat1) > pb->obs_polar) lat1 = copysign(pb->obs_polar, lat1);
    if (fabs(lat2) > pb->obs_polar) lat2 = copysign(pb->obs_polar, lat2);
    if (by1 == by2) return orc_latlon_dist(lat1, glon_at(pb, ys1), lat2, glon_at(pb, ys2));
    double d1 = orc_latlon_dist(lat1, glon_at(pb, ys1), lat2, glon_at(pb, ye2 % pb->nlon + 1));
    double d2 = orc_latlon_dist(lat1, glon_at(pb, ye1 % pb->nlon + 1), lat2, glon_at(pb, ys2));
    return d1 < d2 ? d1 : d2;
}

typedef struct {
    const orc_problem *pb;
    const split_t *sp;
    const int *site_block;    /* block id (bx - 1) * nby + (by - 1) of every site, -1 dropped */
    const double *xb, *axb;
    double *xa, *axa;
    int bx, by;               /* this rank's block (1-based) */
    int sample;               /* every sample-th column of the block */
    long long n_an, n_cand;
    double t_s;
} rank_t;

static void core_blas(int k, int p, const double *hx_in, const float *rdi, const double *dep,
                      const double *rloc, double infl, double *trans, double *w, double *hx,
                      double *work1, double *eivec, double *pa, double *work2, double *eival,
                      double *wk, int lwk, int *iwk, int liwk)
{
    /* letkf_core.F90:58-225 with dgemm/dsyevd */
    double *rr = work2;       /* p entries, reused before work2 is formed */
    for (int j = 0; j < p; ++j) {
        rr[j] = sqrt((double)rdi[j] * rloc[j]);
        for (int i = 0; i < k; ++i) hx[(size_t)j * k + i] = hx_in[(size_t)j * k + i] * rr[j];
    }
    const double one = 1.0, zero = 0.0;
    p_dgemm("n", "t", &k, &k, &p, &one, hx, &k, hx, &k, &zero, work1, &k);
    int info = 0;
    p_dsyevd("V", "L", &k, work1, &k, eival, wk, &lwk, iwk, &liwk, &info);
    for (int j = 0; j < k; ++j) eival[j] = eival[j] + (double)(k - 1) / infl;
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            eivec[j * k + i] = work1[j * k + i];
            work1[j * k + i] /= eival[j];
        }
    p_dgemm("n", "t", &k, &k, &k, &one, work1, &k, eivec, &k, &zero, pa, &k);
    for (int j = 0; j < p; ++j)
        for (int i = 0; i < k; ++i) hx[(size_t)j * k + i] *= rr[j];
    p_dgemm("n", "n", &k, &p, &k, &one, pa, &k, hx, &k, &zero, work2, &k);
    for (int i = 0; i < k; ++i) w[i] = 0.0;
    for (int j = 0; j < p; ++j)
        for (int i = 0; i < k; ++i) w[i] += work2[(size_t)j * k + i] * dep[j];
    for (int j = 0; j < k; ++j) {
        const double r = sqrt((double)(k - 1) / eival[j]);
        for (int i = 0; i < k; ++i) work1[j * k + i] = eivec[j * k + i] * r;
    }
    p_dgemm("n", "t", &k, &k, &k, &one, work1, &k, eivec, &k, &zero, trans, &k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) trans[j * k + i] += w[i];
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void *rank_main(void *arg)
{
    rank_t *r = (rank_t *)arg;
    const orc_problem *pb = r->pb;
    const split_t *sp = r->sp;
    const orc_obs *ob = pb->obs;
    const int k = pb->k, n = pb->n, h = pb->halo;
    double t0 = now_s();
    /* comm_info_init: the neighbour blocks in reach */
    int xs, xe, ys, ye;
    block_range(r->bx, pb->nlat, sp->nbx, &xs, &xe);
    block_range(r->by, pb->nlon, sp->nby, &ys, &ye);
    int north = 0, south = 0;
    for (int b = r->bx - 1; b >= 1; --b) {
        int s2, e2;
        block_range(b, pb->nlat, sp->nbx, &s2, &e2);
        if (orc_latlon_dist(glat_at(pb, xs), glon_at(pb, ys), glat_at(pb, e2 + 1), glon_at(pb, ys)) >
            pb->local_dist) break;
        ++north;
    }
    for (int b = r->bx + 1; b <= sp->nbx; ++b) {
        int s2, e2;
        block_range(b, pb->nlat, sp->nbx, &s2, &e2);
        if (orc_latlon_dist(glat_at(pb, xe + 1), glon_at(pb, ys), glat_at(pb, s2), glon_at(pb, ys)) >
            pb->local_dist) break;
        ++south;
    }
    char *inreach = (char *)calloc((size_t)sp->nbx * sp->nby, 1);
    for (int i = -north; i <= south; ++i) {
        int bx2 = r->bx + i, west = 0, east = 0, by2 = r->by;
        while (west <= (sp->nby - 1) / 2) {
            by2 = (by2 + sp->nby - 2) % sp->nby + 1;
            if (nearest_dist(sp, r->bx, r->by, bx2, by2) > pb->local_dist) break;
            ++west;
        }
        by2 = r->by;
        while (east <= (sp->nby - 1) / 2) {
            by2 = by2 % sp->nby + 1;
            if (nearest_dist(sp, r->bx, r->by, bx2, by2) > pb->local_dist) break;
            ++east;
        }
        if (west + east > sp->nby - 1) --east;
        for (int j = -west; j <= east; ++j) {
            int byj = ((r->by - 1 + j) % sp->nby + sp->nby) % sp->nby + 1;
            inreach[(bx2 - 1) * sp->nby + (byj - 1)] = 1;
        }
    }
    /* the rank's candidate sites */
    int ncand = 0;
    int *cand = (int *)malloc(sizeof(int) * (ob->n_sites > 0 ? ob->n_sites : 1));
    for (int s = 0; s < ob->n_sites; ++s)
        if (r->site_block[s] >= 0 && inreach[r->site_block[s]]) cand[ncand++] = s;
    free(inreach);
    int maxrows = 1;
    for (int c = 0; c < ncand; ++c) {
        int s = cand[c];
        maxrows += ob->site_row_ptr ? ob->site_row_ptr[s + 1] - ob->site_row_ptr[s] : 1;
    }
    double *yb = (double *)malloc(sizeof(double) * (size_t)k * maxrows);
    double *hx = (double *)malloc(sizeof(double) * (size_t)k * maxrows);
    double *work2 = (double *)malloc(sizeof(double) * (size_t)k * maxrows + sizeof(double) * maxrows);
    double *dis = (double *)malloc(sizeof(double) * maxrows);
    float *rdi = (float *)malloc(sizeof(float) * maxrows);
    double *rloc = (double *)malloc(sizeof(double) * maxrows);
    double *work1 = (double *)malloc(sizeof(double) * k * k), *eivec = (double *)malloc(sizeof(double) * k * k);
    double *pa = (double *)malloc(sizeof(double) * k * k), *trans = (double *)malloc(sizeof(double) * k * k);
    double *w = (double *)malloc(sizeof(double) * k), *eival = (double *)malloc(sizeof(double) * k);
    int lwk = 2 * k * k + 6 * k + 1, liwk = 5 * k + 3;
    double *wk = (double *)malloc(sizeof(double) * lwk);
    int *iwk = (int *)malloc(sizeof(int) * liwk);
    double *xa_col = (double *)malloc(sizeof(double) * (size_t)k * n);
    long long col = 0;
    for (int lo = ys; lo <= ye; ++lo)
        for (int la = xs; la <= xe; ++la, ++col) {
            if (col % r->sample) continue;
            const double clat = (double)pb->grid_lat[la - 1], clon = (double)pb->grid_lon[lo - 1];
            int p = 0;
            for (int c = 0; c < ncand; ++c) {
                const int s = cand[c];
                double d = orc_latlon_dist(ob->site_lat[s], ob->site_lon[s], clat, clon);
                double gc = pb->gc_filter ? orc_schr2(d * 1.0 / pb->local_dist)
                                          : orc_schr(d * 1.0 / pb->local_dist);
                if (gc > 0) {
                    const int r0 = ob->site_row_ptr ? ob->site_row_ptr[s] : s;
                    const int r1 = ob->site_row_ptr ? ob->site_row_ptr[s + 1] : s + 1;
                    for (int row = r0; row < r1; ++row, ++p) {
                        memcpy(yb + (size_t)p * k, ob->yb + (size_t)row * k, sizeof(double) * k);
                        dis[p] = ob->dep[row];
                        rdi[p] = (float)ob->rinv[row];
                        rloc[p] = gc;
                    }
                }
            }
            r->n_cand += ncand;
            if (p_dgemm)
                core_blas(k, p, yb, rdi, dis, rloc, pb->inflation_factor, trans, w, hx, work1, eivec,
                          pa, work2, eival, wk, lwk, iwk, liwk);
            else
                orc_letkf_core(k, p, yb, rdi, dis, rloc, trans, w, pb->inflation_factor, 0);
            /* xb_2_xa for the column */
            const size_t off = (((size_t)(lo - 1 + h)) * (pb->nlat + 2 * h) + (la - 1 + h)) * n;
            const double one = 1.0, zero = 0.0;
            if (p_dgemm) {
                p_dgemm("t", "n", &k, &n, &k, &one, trans, &k, r->xb + off * k, &k, &zero, xa_col, &k);
            } else {
                for (int v = 0; v < n; ++v)
                    for (int e = 0; e < k; ++e) {
                        double s = 0.0;
                        for (int f = 0; f < k; ++f) s += r->xb[(off + v) * k + f] * trans[e * k + f];
                        xa_col[v * k + e] = s;
                    }
            }
            for (int v = 0; v < n; ++v) {
                double m = 0.0;
                for (int e = 0; e < k; ++e) {
                    xa_col[v * k + e] += r->axb[off + v];
                    m += xa_col[v * k + e];
                }
                m /= k;
                r->axa[off + v] = m;
                for (int e = 0; e < k; ++e) r->xa[(off + v) * k + e] = xa_col[v * k + e] - m;
            }
            r->n_an += 1;
        }
    free(cand); free(yb); free(hx); free(work2); free(dis); free(rdi); free(rloc);
    free(work1); free(eivec); free(pa); free(trans); free(w); free(eival); free(wk); free(iwk);
    free(xa_col);
    r->t_s = now_s() - t0;
    return NULL;
}

/* index of the last grid point <= x (0 when x is below the first) */
static int last_le(const float *g, int n, double x)
{
    int lo = 0, hi = n - 1;
    if (n <= 0 || (double)g[0] > x) return 0;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if ((double)g[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/*
 * Runs the horizontal loop (cube_n = 1) of pb with nthreads ranks on every
 * sample-th column of each rank's block.  xa/axa receive the analysed
 * columns.  Outputs: analyses done, total candidate sites scanned, wall time
 * of the slowest rank (the step's time on P ranks), the split.
 */
int orc_baseline_run(const orc_problem *pb, const double *xb, const double *axb, double *xa,
                     double *axa, int nthreads, int sample, long long *n_an, long long *n_cand,
                     double *wall_s, int *split)
{
    pthread_once(&g_once, load_blas);
    if (pb->vert_dist > 0.0 || (pb->cube_n > 1) || nthreads < 1 || sample < 1) return -1;
    split_t sp;
    sp.pb = pb;
    int d0, d1;
    dims_create(nthreads, &d0, &d1);
    /* dims(1) splits longitude, dims(2) latitude (module_da_parallel.F90:84-87) */
    sp.nby = d0;
    sp.nbx = d1;
    if (sp.nby > pb->nlon) sp.nby = pb->nlon;
    if (sp.nbx > pb->nlat) sp.nbx = pb->nlat;
    const orc_obs *ob = pb->obs;
    int *site_block = (int *)malloc(sizeof(int) * (ob->n_sites > 0 ? ob->n_sites : 1));
    /* the rank whose block holds the site's grid cell reads it */
    for (int s = 0; s < ob->n_sites; ++s) {
        if (fabs(ob->site_lat[s]) > pb->obs_polar) { site_block[s] = -1; continue; }
        const int ila = last_le(pb->grid_lat, pb->nlat, ob->site_lat[s]);
        const int ilo = last_le(pb->grid_lon, pb->nlon, ob->site_lon[s]);
        int bx = 1, by = 1, a, b;
        for (bx = 1; bx <= sp.nbx; ++bx) { block_range(bx, pb->nlat, sp.nbx, &a, &b); if (ila + 1 <= b) break; }
        for (by = 1; by <= sp.nby; ++by) { block_range(by, pb->nlon, sp.nby, &a, &b); if (ilo + 1 <= b) break; }
        site_block[s] = (bx - 1) * sp.nby + (by - 1);
    }
    const int P = sp.nbx * sp.nby;
    rank_t *rk = (rank_t *)calloc(P, sizeof(rank_t));
    pthread_t *th = (pthread_t *)calloc(P, sizeof(pthread_t));
    double t0 = now_s();
    for (int i = 0; i < P; ++i) {
        rk[i].pb = pb; rk[i].sp = &sp; rk[i].site_block = site_block;
        rk[i].xb = xb; rk[i].axb = axb; rk[i].xa = xa; rk[i].axa = axa;
        rk[i].bx = i / sp.nby + 1; rk[i].by = i % sp.nby + 1; rk[i].sample = sample;
        pthread_create(&th[i], NULL, rank_main, &rk[i]);
    }
    long long na = 0, nc = 0;
    for (int i = 0; i < P; ++i) {
        pthread_join(th[i], NULL);
        na += rk[i].n_an;
        nc += rk[i].n_cand;
    }
    *wall_s = now_s() - t0;
    *n_an = na;
    *n_cand = nc;
    split[0] = sp.nbx;
    split[1] = sp.nby;
    free(rk); free(th); free(site_block);
    return 0;
}
