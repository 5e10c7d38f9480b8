/*
 * letkf_oracle.h -- CPU restatement of the reference LETKF path.
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline; never the product).
 */
#ifndef LETKF_ORACLE_H
#define LETKF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Observation set: sites (lat, lon) each owning a contiguous run of rows
 * (level x var, level-major as letkf_yb.F90:102-103).  site_row_ptr == NULL
 * means one row per site. */
typedef struct {
    int n_sites;
    int n_rows;
    const double *site_lat;      /* degrees */
    const double *site_lon;      /* degrees */
    const int *site_row_ptr;     /* n_sites + 1, or NULL */
    const double *yb;            /* n_rows x k, member index fastest */
    const double *dep;           /* n_rows: y - mean H(x) */
    const double *rinv;          /* n_rows: 1/sigma^2 (cast to real(4) on use) */
    const double *vcoord;        /* n_rows: level (LEVEL mode) or ln p (LNP mode); may be NULL */
} orc_obs;

enum { ORC_VCOORD_LEVEL = 0, ORC_VCOORD_LNP = 1 };
enum { ORC_EIG_JACOBI = 1, ORC_GUARD = 2 };

typedef struct {
    int k;                 /* ensemble size */
    int n;                 /* state elements per column */
    int nlat, nlon;        /* interior grid (reference nx = lat, ny = lon) */
    int halo;              /* halo width of xb/xa arrays */
    int nz;                /* levels of the 3-D variables */
    int cube_n;            /* da_block_interval */
    int cube_nz;           /* da_vertical_interval */
    const float *grid_lat; /* nlat, real(4) degrees */
    const float *grid_lon; /* nlon, real(4) degrees */
    double local_dist;     /* da_loc_distance (km); GC support is 2*local_dist */
    double vert_dist;      /* da_vert_distance; <= 0 selects the horizontal loop */
    double inflation_factor;
    double obs_polar;      /* obs with |lat| > obs_polar are skipped */
    int gc_filter;         /* schr2 (1) or schr boxcar (0) */
    int gc_filter_vertical;
    int max_obs;           /* <= 0: no cap */
    int core_flags;        /* ORC_EIG_JACOBI | ORC_GUARD */
    int vcoord_mode;       /* ORC_VCOORD_LEVEL / ORC_VCOORD_LNP */
    int ps_index;          /* 0-based state index of ps (LNP mode) */
    int n_var;             /* number of DA variables */
    const int *var_start;  /* 1-based start in n (da_var_start) */
    const int *var_lev;    /* 1 for 2-D vars, nz for 3-D vars (da_var_lev) */
    const double *hyai, *hybi; /* nz+1 hybrid coefficients (LNP mode) */
    double p0, ptop;
    const orc_obs *obs;
    int overlap;           /* da_block_overlap: cubes every cube_n - overlap columns
                              (letkf_compute, horizontal loop only); 0: no overlap */
} orc_problem;

double orc_schr2(double ri);
double orc_schr(double ri);
double orc_pi_f32(void);
double orc_earth_radius_km(void);
double orc_latlon_dist(double x1, double y1, double x2, double y2);
int orc_gen_yb_dis(const orc_obs *obs, int k, double center_lat, double center_lon,
                   double local_dist, int gc_filter, double obs_polar, int max_obs,
                   int cap, double *yb, double *dis, float *rdiaginv, double *rloc,
                   double *vco);
void orc_vertical_localization(const double *rloc_in, double *rloc, int nobsl,
                               double lnp, const double *obs_lnp, double vert_dist,
                               int gc_filter_vertical);
int orc_syev_ql(int n, const double *a, double *w, double *v);
int orc_syev_jacobi(int n, const double *a, double *w, double *v);
int orc_letkf_core(int nanals, int nobsl, const double *hxens_in, const float *rdiaginv,
                   const double *dep, const double *rloc, double *trans, double *overline_w,
                   double inflation_factor, int flags);
int orc_letkf_transform(int k, const double *a, const double *b, double inflation_factor,
                        int flags, double *trans, double *overline_w);
void orc_xb_2_xa(int k, int m, const double *xb, const double *axb, double *xa, double *axa,
                 const double *trans);
int orc_n_analyses(const orc_problem *pb);
int orc_letkf_compute(const orc_problem *pb, const double *xb, const double *axb, double *xa,
                      double *axa, double *trans_out, double *w_out, int *nobsl_out,
                      int *status_out, const int *h_list, int n_list, int nthreads);
void orc_epilogue(const orc_problem *pb, const double *xb, const double *axb, const double *xa,
                  const double *axa, double polar_lat, double gamma1, double gamma2,
                  double *xa_out, double *spread_b, double *spread_a);

/* Observation operator and Yb/d centring of gen_obs_struct.
 * Rows of one site are level-major (all variables of a level together) and
 * carry the level pressure row_ph (Pa) and the 0-based DA variable row_var.
 * grid_lat/grid_lon of pb are closed by 90 and 360 (pro_info.F90:77-78);
 * xb/axb need a halo >= 1.  yo is the observed value; outputs are
 * yb (n_rows x k, member fastest) and dep = yo - mean_e H(x_e).
 * Returns 0, or -1 for bad arguments. */
int orc_obs_operator(const orc_problem *pb, const double *xb, const double *axb, int n_sites,
                     const double *site_lat, const double *site_lon, const int *site_row_ptr,
                     const int *row_var, const double *row_ph, const double *yo, double *yb,
                     double *dep);

/* CPU baseline (letkf_baseline.c): the horizontal loop on nthreads ranks of
 * the reference's block split, each scanning its comm_info_init candidate
 * sites, letkf_core/xb_2_xa on MKL dgemm/dsyevd when MKL loads.  Every
 * sample-th column of each block.  Returns 0, or -1 for an unsupported
 * problem (vertical loop, cube_n > 1). */
int orc_baseline_run(const orc_problem *pb, const double *xb, const double *axb, double *xa,
                     double *axa, int nthreads, int sample, long long *n_an, long long *n_cand,
                     double *wall_s, int *split);
const char *orc_baseline_provider(void);

#ifdef __cplusplus
}
#endif
#endif
