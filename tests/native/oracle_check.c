/*
 * The CPU oracle and baseline under -fsanitize=address,undefined (built and
 * run by tests/test_sanitize.py): the analysis loops (horizontal, vertical,
 * overlapped cubes), the epilogue and the CPU baseline on small generated
 * problems, with checks that the outputs are finite and that R^-1 = 0 leaves
 * the state unchanged.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/letkf_oracle.h"

static unsigned long long rs = 88172645463325252ULL;
static double urand(void)
{
    rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
    return (double)(rs >> 11) / 9007199254740992.0;
}

static int check_finite(const char *what, const double *a, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(a[i])) { printf("FAIL %s not finite at %zu\n", what, i); return 1; }
    return 0;
}

int main(void)
{
    enum { K = 8, NLAT = 7, NLON = 12, NZ = 3, HALO = 1, NSITE = 90 };
    const int n = 1 + 2 * NZ;                 /* ps + two 3-D variables */
    float glat[NLAT], glon[NLON];
    for (int i = 0; i < NLAT; ++i) glat[i] = (float)(-60.0 + 20.0 * i);
    for (int i = 0; i < NLON; ++i) glon[i] = (float)(30.0 * i);
    double slat[NSITE], slon[NSITE], yb[NSITE * K], dep[NSITE], rinv[NSITE], vco[NSITE];
    int rp[NSITE + 1];
    for (int s = 0; s < NSITE; ++s) {
        slat[s] = -70.0 + 140.0 * urand();
        slon[s] = 360.0 * urand();
        rp[s] = s;
        double m = 0.0;
        for (int e = 0; e < K; ++e) { yb[s * K + e] = urand() - 0.5; m += yb[s * K + e]; }
        for (int e = 0; e < K; ++e) yb[s * K + e] -= m / K;
        dep[s] = urand() - 0.5;
        rinv[s] = 1.0 + urand();
        vco[s] = 1.0 + (NZ - 1) * urand();
    }
    rp[NSITE] = NSITE;
    orc_obs obs = {NSITE, NSITE, slat, slon, rp, yb, dep, rinv, vco};
    int var_start[3] = {1, 2, 2 + NZ}, var_lev[3] = {1, NZ, NZ};
    orc_problem pb;
    memset(&pb, 0, sizeof(pb));
    pb.k = K; pb.n = n; pb.nlat = NLAT; pb.nlon = NLON; pb.halo = HALO; pb.nz = NZ;
    pb.cube_n = 1; pb.cube_nz = 2; pb.grid_lat = glat; pb.grid_lon = glon;
    pb.local_dist = 1500.0; pb.vert_dist = 0.0; pb.inflation_factor = 1.1; pb.obs_polar = 91.0;
    pb.gc_filter = 1; pb.gc_filter_vertical = 1; pb.max_obs = 0; pb.core_flags = 0;
    pb.vcoord_mode = 0; pb.ps_index = 0; pb.n_var = 3; pb.var_start = var_start;
    pb.var_lev = var_lev; pb.obs = &obs; pb.overlap = 0;
    const size_t cols = (size_t)(NLAT + 2 * HALO) * (NLON + 2 * HALO);
    double *xb = malloc(sizeof(double) * cols * n * K), *axb = malloc(sizeof(double) * cols * n);
    double *xa = calloc(cols * n * K, sizeof(double)), *axa = calloc(cols * n, sizeof(double));
    for (size_t c = 0; c < cols * n; ++c) {
        double m = 0.0;
        for (int e = 0; e < K; ++e) { xb[c * K + e] = urand() - 0.5; m += xb[c * K + e]; }
        for (int e = 0; e < K; ++e) xb[c * K + e] -= m / K;
        axb[c] = 10.0 * urand();
    }
    int bad = 0;
    /* horizontal loop, cube_n = 1 and 3, then the overlapped cubes */
    const int cubes[3][2] = {{1, 0}, {3, 0}, {4, 2}};
    for (int t = 0; t < 3; ++t) {
        pb.cube_n = cubes[t][0];
        pb.overlap = cubes[t][1];
        int na = orc_n_analyses(&pb);
        double *tr = malloc(sizeof(double) * na * K * K), *w = malloc(sizeof(double) * na * K);
        int *nob = malloc(sizeof(int) * na), *st = malloc(sizeof(int) * na);
        if (orc_letkf_compute(&pb, xb, axb, xa, axa, tr, w, nob, st, NULL, 0, 3)) {
            printf("FAIL letkf_compute cube_n=%d overlap=%d\n", pb.cube_n, pb.overlap);
            bad = 1;
        }
        bad |= check_finite("xa", xa, cols * n * K) | check_finite("trans", tr, (size_t)na * K * K);
        free(tr); free(w); free(nob); free(st);
    }
    pb.cube_n = 1;
    pb.overlap = 0;
    /* vertical loop */
    pb.vert_dist = 1.5;
    if (orc_letkf_compute(&pb, xb, axb, xa, axa, NULL, NULL, NULL, NULL, NULL, 0, 2)) {
        printf("FAIL vertical loop\n");
        bad = 1;
    }
    bad |= check_finite("xa vertical", xa, cols * n * K);
    pb.vert_dist = 0.0;
    /* epilogue */
    double *xo = malloc(sizeof(double) * cols * n * K);
    double *sb = malloc(sizeof(double) * NLAT * NLON * n), *sa = malloc(sizeof(double) * NLAT * NLON * n);
    orc_epilogue(&pb, xb, axb, xa, axa, 50.0, 0.3, 0.7, xo, sb, sa);
    bad |= check_finite("xa_out", xo, cols * n * K) | check_finite("spread", sa, NLAT * NLON * n);
    /* R^-1 = 0 leaves the interior unchanged */
    for (int s = 0; s < NSITE; ++s) rinv[s] = 0.0;
    pb.inflation_factor = 1.0;
    orc_letkf_compute(&pb, xb, axb, xa, axa, NULL, NULL, NULL, NULL, NULL, 0, 1);
    for (int lo = 0; lo < NLON; ++lo)
        for (int la = 0; la < NLAT; ++la) {
            size_t c = ((size_t)(lo + HALO) * (NLAT + 2 * HALO) + la + HALO) * n;
            for (int v = 0; v < n; ++v) {
                if (fabs(axa[c + v] - axb[c + v]) > 1e-12) { printf("FAIL axa %d %d\n", lo, la); bad = 1; }
                for (int e = 0; e < K; ++e)
                    if (fabs(xa[(c + v) * K + e] - xb[(c + v) * K + e]) > 1e-12) {
                        printf("FAIL xa %d %d\n", lo, la);
                        bad = 1;
                    }
            }
        }
    /* the CPU baseline (the reference's rank blocks), every column */
    for (int s = 0; s < NSITE; ++s) rinv[s] = 1.0;
    long long nan_ = 0, ncand = 0;
    double wall = 0.0;
    int split[2] = {0, 0};
    if (orc_baseline_run(&pb, xb, axb, xa, axa, 4, 1, &nan_, &ncand, &wall, split) != 0 || nan_ <= 0) {
        printf("FAIL baseline\n");
        bad = 1;
    }
    free(xb); free(axb); free(xa); free(axa); free(xo); free(sb); free(sa);
    if (!bad) printf("oracle_check ok (baseline on %s)\n", orc_baseline_provider());
    return bad;
}
