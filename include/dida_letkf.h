/*
 * dida_letkf.h -- C ABI of the MI355X LETKF analysis engine.
 *
 * Drop-in boundary for the x86 DA LETKF analysis step of the reference
 * (Fortran, x86/DA/LETKF/src).  Every entry point takes plain pointers and
 * sizes; arrays use the reference's Fortran (column-major) layouts:
 *   xb(ens, n, 1-halo:nlat+halo, 1-halo:nlon+halo)   (letkf_compute.F90:15)
 *   axb(n, 1-halo:nlat+halo, 1-halo:nlon+halo)        (letkf_compute.F90:16)
 *   obs rows pro_yb(ens, .) with the member index fastest (read_obs_mod.F90:25)
 *   trans(ens, ens), overline_w(ens)                  (letkf_core.F90:26-27)
 * The reference's nx is the latitude count and ny the longitude count
 * (letkf.F90:21); here they are called nlat and nlon.
 *
 * Ownership: the caller owns every array it passes; the library copies what
 * it needs and keeps no caller pointer after a call returns.  Device-pointer
 * variants (*_device) take memory on the current HIP device and run on the
 * handle's stream.
 *
 * Errors: functions return DIDA_OK (0), a negative DIDA_ERR_* code, or, from
 * the analysis calls, a positive count of analyses whose eigensolve failed
 * (the reference prints dsyevd's ierr and continues, letkf_core.F90:119-122);
 * those analyses leave the background in place and give NaN transforms.
 * dida_letkf_last_error() gives the message of the last failure on the
 * calling thread.  No call aborts the process.
 *
 * Threading: a handle is used by one host thread at a time; different
 * handles are independent.
 */
#ifndef DIDA_LETKF_H
#define DIDA_LETKF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIDA_LETKF_ABI_VERSION 1

enum {
    DIDA_OK = 0,
    DIDA_ERR_ARG = -1,          /* invalid argument */
    DIDA_ERR_HIP = -2,          /* HIP runtime error */
    DIDA_ERR_NOMEM = -3,        /* device allocation failed */
    DIDA_ERR_STATE = -4,        /* call out of order (no grid / no obs) */
    DIDA_ERR_ADMISSION = -5,    /* serving: job refused by the HBM budget */
    DIDA_ERR_UNSUPPORTED = -6   /* configuration not supported */
};

/* Precision modes of the obs-space products.  FP64 is the reference
 * letkf_core (computing_precision >= 52).  The reduced modes follow
 * letkf_core_mpal's pattern (letkf_core_mpal.F90): low-precision obs-space
 * products, FP64 eigendecomposition and inverse square root, eigenvalue
 * guard.  The ensemble update runs in FP64 on the double-precision state. */
enum {
    DIDA_PREC_FP64 = 0,     /* FP64 products (MFMA f64) */
    DIDA_PREC_FP32 = 1,     /* FP32 MFMA products, FP64 eigensolve */
    DIDA_PREC_BF16X3 = 2    /* BF16 hi/lo split MFMA products (3 terms, FP32
                               accumulation), FP64 eigensolve */
};

/* Vertical coordinate of the vertical-localisation loop. */
enum {
    DIDA_VCOORD_LEVEL = 0,  /* level index (obs vcoord in level units) */
    DIDA_VCOORD_LNP = 1     /* ln p: obs vcoord = log(ph) (letkf_yb.F90:211),
                               column profile from the hybrid p(ps) */
};

/* Mirrors namelist_da / letkf_init globals (da_namelist_mod.F90:27-49,
 * letkf_init.F90:27-81). */
typedef struct dida_letkf_config {
    int k;                  /* ensemble size (svr_ensemble), 2..128 */
    int nlat, nlon;         /* interior grid of the block (nx, ny) */
    int nz;                 /* levels of 3-D variables (num_lev) */
    int n2d, n3d;           /* 2-D vars first, then nz levels per 3-D var:
                               n = n2d + n3d*nz (letkf_init.F90:63-73) */
    int halo;               /* da_halo of xb/xa/axb/axa */
    int cube_n;             /* da_block_interval (horizontal cube) */
    int cube_nz;            /* da_vertical_interval (vertical chunk) */
    double local_dist;      /* da_loc_distance, km (GC support 2*local_dist) */
    double vert_dist;       /* da_vert_distance; <= 0 runs the horizontal loop */
    double inflation_factor;/* da_inflation_factor */
    double obs_polar;       /* obs with |lat| > obs_polar are skipped */
    int gc_filter;          /* da_gc_enable: schr2 (1) or boxcar schr (0) */
    int gc_filter_vertical; /* da_gc_enable_vertical */
    int precision;          /* DIDA_PREC_* */
    int vcoord_mode;        /* DIDA_VCOORD_* */
    int ps_index;           /* 0-based index of ps in n (ps_pos - 1) */
    const double *hyai;     /* nz+1 hybrid coefficients (DIDA_VCOORD_LNP) */
    const double *hybi;
    double p0, ptop;
    int keep_transforms;    /* keep trans/overline_w of every analysis */
    int max_batch_columns;  /* 0 = automatic */
    int transform_kernel;   /* 0: split eigensolver (tridiagonalisation,
                               lane-per-matrix QL, rotation apply), a batch
                               of 2048 analyses or more as two half-batches
                               on two streams; 1: the single-kernel
                               transform (k <= 64); 2: the split eigensolver
                               on the caller's stream only (its kernels
                               timed one at a time) */
    int products_kernel;    /* 0: automatic (= 1); 1: one selection list per
                               analysis; 2: tiles of 8 analyses sharing their
                               LDS-staged union of obs rows (horizontal loop,
                               k <= 64; other loops use 1) */
    int overlap;            /* da_block_overlap: 0 runs the no-overlap loops;
                               1 .. cube_n-1 runs letkf_compute
                               (letkf_compute.F90:156-256): cubes of cube_n
                               columns every cube_n - overlap columns, each
                               column the mean of its cubes' analyses
                               (horizontal loop, FP64 state) */
} dida_letkf_config;

typedef struct dida_letkf_stats {
    long long n_analyses;   /* local analyses (k x k transform solves) */
    long long total_rows;   /* sum of local obs rows over analyses */
    int max_rows;           /* largest local obs count */
    int n_flagged;          /* analyses whose eigensolve failed on both passes (their
                               state keeps the background: xa = xb, axa = axb) */
    int rows_over_10000;    /* analyses the reference's MAX_OBS would truncate */
    double ms_select;       /* device time, local obs selection */
    double ms_products;     /* device time, obs-space products */
    double ms_transform;    /* device time, eigensolve + transform + update */
    double ms_total;        /* device time of the whole call */
    double ms_tridiag;      /* of ms_transform: Householder tridiagonalisation kernel */
    double ms_ql;           /* of ms_transform: QL recurrence kernel */
    double ms_apply;        /* of ms_transform: eigenvector and transform kernel (k <= 64
                               with the fused kernel: also the update) */
    double ms_update;       /* of ms_transform: ensemble update kernel (0 when fused) */
    int n_rerun;            /* analyses the eigensolver's first pass flagged (QL not
                               converged in 30 iterations, or its rotation log full)
                               and its list pass redid */
    /* Products schedule.  Length buckets over the local obs count p of a
     * column: [0] p = 0, [1] 1..256, [2] 257..512, ... [6] 4097..8192, [7] more. */
    int schedule;                    /* 1: buckets + persistent work queue; 0: column order */
    long long bucket_analyses[8];    /* analyses per bucket */
    long long bucket_rows[8];        /* their local obs rows */
    double bucket_wave_ms[8];        /* DIDA_OPT_TIMING: products time per bucket summed
                                        over the queue's workers (worker-ms) */
    double products_balance;         /* DIDA_OPT_TIMING: busy fraction of the products
                                        workers over their launch spans (1: none idle) */
    int variant[8];                  /* products variant per bucket (32-row groups per step) */
    int graph_replay;                /* 1: the call ran as a hipGraph launch (ms_total is the
                                        graph's time; the per-phase times are 0) */
    int reran_rows;                  /* 1: the row lists outgrew their capacity and the call
                                        ran again with a larger one */
    long long row_cap;               /* row-list capacity per batch after the call */
} dida_letkf_stats;

typedef struct dida_letkf_handle dida_letkf_handle;

/* Defaults of namelist_da where they exist (cube_n = 1 instead of 4,
 * overlap = 0 instead of 2). */
void dida_letkf_config_default(dida_letkf_config *cfg);

/* Replaces letkf_init (letkf_init.F90:2): validates the configuration and
 * creates device state on the current HIP device. */
int dida_letkf_create(const dida_letkf_config *cfg, dida_letkf_handle **out);
int dida_letkf_destroy(dida_letkf_handle *h);

/* Runs all later device work of the handle on this hipStream_t (NULL: the
 * handle's own stream). */
int dida_letkf_set_stream(dida_letkf_handle *h, void *hip_stream);

/* Scheduling options of a handle.  No counterpart in the reference, whose
 * analyses run one after another per MPI rank (letkf_compute.F90:438-552).
 *   DIDA_OPT_SCHEDULE  1 (default): the products of each batch run over
 *                      length buckets of p, longest first, as persistent
 *                      launches that take analyses from a work queue (one
 *                      launch per run of buckets sharing a variant); 0: one
 *                      launch in column order.  Results are the same bits.
 *   DIDA_OPT_TIMING    1: the queue's kernels time each bucket and their
 *                      workers' busy spans (stats bucket_wave_ms,
 *                      products_balance); 0 (default): off.
 *   DIDA_OPT_TUNE      1: the next eager call measures every products
 *                      variant on each non-empty bucket, keeps the fastest and
 *                      writes them to the tuning cache ($DIDA_TUNE_FILE, else
 *                      $XDG_CACHE_HOME or $HOME/.cache, dida_letkf/tune.tsv),
 *                      keyed by (device, k, precision, bucket); handles read
 *                      the cache when created.  Variants give the same bits.
 *   DIDA_OPT_GRAPH     1: a call captures its launches in a hipGraph (one
 *                      stream, no events) and launches it; later calls with
 *                      the same buffers, sizes, options and obs replay it.
 *                      Needs a non-null stream.  0 (default): eager launches.
 *   DIDA_OPT_ROW_CAP   row-list capacity per batch (0: estimated from the obs
 *                      density).  A call whose lists need more runs again with
 *                      the measured need; the capacity never shrinks by itself.
 * Returns DIDA_ERR_ARG for an unknown option or value. */
enum {
    DIDA_OPT_SCHEDULE = 1,
    DIDA_OPT_TIMING = 2,
    DIDA_OPT_TUNE = 3,
    DIDA_OPT_GRAPH = 4,
    DIDA_OPT_ROW_CAP = 5
};
int dida_letkf_set_option(dida_letkf_handle *h, int option, long long value);

/* Block grid as real(4) degrees, like pro_info's grid_lat/grid_lon
 * (pro_info.F90:11-12); grid_lat has nlat entries, grid_lon nlon. */
int dida_letkf_set_grid(dida_letkf_handle *h, const float *grid_lat, const float *grid_lon);

/* Observation set of one cycle, replacing the obs lists that gen_obs_struct
 * and comm_obs_data build (letkf_obs.F90:16-187,251-466).  Sites (lat, lon
 * in degrees) own contiguous runs of rows; site_row_ptr (n_sites+1 entries,
 * NULL = one row per site) gives each site's rows in level-major, var-minor
 * order (letkf_yb.F90:102-103).  Per row: yb (k values, centred
 * perturbations), dep = y - mean H(x), rinv = 1/sigma^2 (used as real(4)
 * like rdiaginv), vcoord = level or log(ph) (NULL if no vertical loop).
 * Sites with |lat| > obs_polar are dropped here (letkf_yb.F90:87). */
int dida_letkf_set_obs(dida_letkf_handle *h, int n_sites, const double *site_lat,
                       const double *site_lon, const int *site_row_ptr, int n_rows,
                       const double *yb, const double *dep, const double *rinv,
                       const double *vcoord);
int dida_letkf_set_obs_device(dida_letkf_handle *h, int n_sites, const double *site_lat,
                              const double *site_lon, const int *site_row_ptr, int n_rows,
                              const double *yb, const double *dep, const double *rinv,
                              const double *vcoord);

/* The analysis loop of one cycle: letkf_compute_nooverlap
 * (letkf_compute.F90:12-154) when vert_dist <= 0, else
 * letkf_compute_nooverlap_vertical (letkf_compute.F90:386-554).  Host
 * arrays; interior columns of xa/axa are written, halo entries are left as
 * they are.  stats may be NULL. */
int dida_letkf_analyze(dida_letkf_handle *h, const double *xb, const double *axb, double *xa,
                       double *axa, dida_letkf_stats *stats);
/* Same with device-resident arrays (the serving and benchmark path). */
int dida_letkf_analyze_device(dida_letkf_handle *h, const double *xb, const double *axb,
                              double *xa, double *axa, dida_letkf_stats *stats);
/* Same with the perturbations stored in FP32 (xb, xa: float, member fastest,
 * the xb layout) and the means in FP64: the reduced-precision state the
 * reference's xb_type may be (xb_2_xa's sgemm branch, letkf_compute.F90:293).
 * The ensemble update runs on f32 MFMA with FP32 accumulation; the means,
 * the re-centring and the eigensolver stay FP64. */
int dida_letkf_analyze_device_f32(dida_letkf_handle *h, const float *xb, const double *axb,
                                  float *xa, double *axa, dida_letkf_stats *stats);
/* The analysis with the caller's epilogue (run_da_parallel,
 * module_da_parallel.F90:219-261) fused into the update's output pass, on
 * device arrays: instead of xa/axa it writes xa_out (full members, the xb
 * layout, FP64) = gamma1 (axb + xb) + gamma2 (axa + xa) on interior columns
 * with |grid_lat| <= polar_lat, xb + axb on the others, 0 on halo columns,
 * and spread_b/spread_a (n, nlat, nlon) = RMS over members (divisor ens) of
 * xb and of xa_out - axa (xb where |grid_lat| > polar_lat).  The _f32 form
 * reads FP32 perturbations as dida_letkf_analyze_device_f32. */
int dida_letkf_analyze_epilogue_device(dida_letkf_handle *h, const double *xb, const double *axb,
                                       double polar_lat, double gamma1, double gamma2,
                                       double *xa_out, double *spread_b, double *spread_a,
                                       dida_letkf_stats *stats);
int dida_letkf_analyze_epilogue_device_f32(dida_letkf_handle *h, const float *xb,
                                           const double *axb, double polar_lat, double gamma1,
                                           double gamma2, double *xa_out, double *spread_b,
                                           double *spread_a, dida_letkf_stats *stats);

/* Replaces run_da_parallel's dnn2letkf block (module_da_parallel.F90:
 * 186-211) on the device: x_in is the full ensemble state xb_in(ens, n,
 * 1-halo:nlat+halo, 1-halo:nlon+halo); axb = sum_e x_in / ens and
 * xb = x_in - axb for every column, halo included, into xb (FP64) or xb32
 * (FP32; give exactly one).  The reference stops on a NaN member; this call
 * returns DIDA_ERR_ARG naming the first such element (k, j, i). */
int dida_letkf_split_device(dida_letkf_handle *h, const double *x_in, double *xb, float *xb32,
                            double *axb);

/* run_da_parallel's analysis step on the device (module_da_parallel.F90:
 * 186-261): the split above (into buffers of the handle; FP32 perturbations
 * when state_f32), the analysis, and the letkf2dnn epilogue fused into the
 * update as dida_letkf_analyze_epilogue_device does.  One readback at the
 * end; a NaN background member fails the call as above (outputs are then
 * undefined). */
int dida_letkf_run_da_device(dida_letkf_handle *h, const double *x_in, int state_f32,
                             double polar_lat, double gamma1, double gamma2, double *xa_out,
                             double *spread_b, double *spread_a, dida_letkf_stats *stats);

/* Per-analysis outputs of the last analyze call when keep_transforms was
 * set: trans (k*k per analysis, column-major), overline_w (k), nobsl.
 * Any pointer may be NULL.  n_analyses must equal the analysis count. */
int dida_letkf_get_transforms(dida_letkf_handle *h, long long n_analyses, double *trans,
                              double *overline_w, int *nobsl);

/* Number of local analyses the configuration performs per cycle. */
long long dida_letkf_n_analyses(const dida_letkf_config *cfg);

/* Batched letkf_core: n_analyses independent calls of letkf_core
 * (letkf_core.F90:14) whose inputs are packed back to back.  Analysis a uses
 * obs columns row_ptr[a] .. row_ptr[a+1]-1 of hxens (nanals x total, column
 * major), rdiaginv, dep and rloc.  Outputs trans (nanals*nanals per
 * analysis) and overline_w (nanals per analysis); status (may be NULL) gets
 * the eigensolver flag per analysis.  Host arrays. */
int dida_letkf_core_batched(int nanals, long long n_analyses, const long long *row_ptr,
                            const double *hxens, const float *rdiaginv, const double *dep,
                            const double *rloc, double *trans, double *overline_w,
                            double inflation_factor, int precision, int *status);

/* dida_letkf_core_batched without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as
 * dida_letkf_core_batched; every device buffer (the staged inputs, the
 * products, the eigensolver's scratch, the outputs) is carved from one fixed
 * 64 MiB array that each device running the library holds until the process
 * ends.  Calls needing more scratch than that return DIDA_ERR_ARG with the
 * size in the message.  The work runs on the null stream and the call returns
 * when it is done.  Calls on one device share the array: within a process
 * they run one after another (a lock per device); each call starts from
 * cleared state, nothing of an earlier call is read. */
int dida_letkf_core_batched_noalloc(int nanals, long long n_analyses, const long long *row_ptr,
                                    const double *hxens, const float *rdiaginv,
                                    const double *dep, const double *rloc, double *trans,
                                    double *overline_w, double inflation_factor, int precision,
                                    int *status);

/* Drop-in for the external letkf_core (letkf_core.F90:14-33), the same
 * by-reference signature the Sunway C kernel uses
 * (sw/DA/LETKF/src/letkf_half/letkf_core_half_master.c:24).  Runs one
 * analysis on the GPU in FP64; pflag is accepted and ignored.  Each calling
 * thread keeps, per device, a stream and the device buffers of its largest
 * call until the process ends, so a call costs its copies and launches.  On
 * an error, trans and overline_w are NaN and the error is printed to stderr
 * (the reference's interface has no status). */
void letkf_core_(const int *nanals, const int *nobsl, const double *hxens_in,
                 const float *rdiaginv, const double *dep, const double *rloc, double *trans,
                 double *overline_w, const double *inflation_factor, const int *pflag);

/* Observation records as read_obs receives them from its store
 * (read_obs_mod.F90:338-449; RedisGetobserveF's buffers), host arrays:
 * record i has lon[i], lat[i], lev[i] (real(4); lev is the 1-based level of a
 * multi-level type) and val[(2*nvar+1)*i ...] = value_1, oe_1, ..., value_nvar,
 * oe_nvar, ph.  multi_level != 0 (ADPUPA, GPSRO): a record with lev == 1 starts
 * a site and the records after it are its other levels (:411-428); otherwise
 * every record is a site of one level (:429-437).
 * dida_obs_records_count gives the sites and rows (nvar per record);
 * dida_obs_records_convert fills the inputs of dida_letkf_obs_operator_device
 * and dida_letkf_set_obs[_device]: per site lat/lon and site_row_ptr
 * (n_sites + 1), per row (level-major, the record's variables in order)
 * row_var = var_index[n] (the DA variable of obs variable n, which read_obs
 * matches by name), row_ph = ph, yo = value + noise * oe and
 * rinv = 1 / oe**2, both in real(4) as the reference computes them.  noise
 * (nvar per record, standard normal draws; NULL: none) replaces the
 * reference's randn.  Both return DIDA_ERR_ARG for a profile record before
 * any lev == 1 (which the reference would store at site 0) or a level that
 * is not a positive whole number.  No device work. */
int dida_obs_records_count(int n_records, int nvar, int multi_level, const float *lev,
                           int *n_sites, int *n_rows);
int dida_obs_records_convert(int n_records, int nvar, int multi_level, const float *lon,
                             const float *lat, const float *lev, const float *val,
                             const float *noise, const int *var_index, double *site_lat,
                             double *site_lon, int *site_row_ptr, int *row_var, double *row_ph,
                             double *yo, double *rinv);

/* Observation operator and Yb/d centring of gen_obs_struct
 * (letkf_obs.F90:16-187: find_obs_grid :192-248, ps2ph_3d_interp and
 * obs_interpolation_3d_2 obs_make.F90:233-381, centring :158-170) on device
 * arrays, on the handle's stream.  Needs set_grid, halo >= 1, hyai/hybi
 * (nz + 1 hybrid interface coefficients) in the config, and ps at state index
 * ps_index.  Sites: lat/lon in degrees and site_row_ptr (n_sites + 1); the
 * rows of a site are level-major (all variables of one level together, levels
 * in the order of the obs profile) with row_var (0-based DA variable) and
 * row_ph (level pressure, Pa); yo holds the observed values.  Writes
 * yb (n_rows x k, member fastest) = H(x_e) - mean_e H(x_e) and
 * dep (n_rows) = yo - mean_e H(x_e), the inputs of set_obs_device.  A row
 * whose variable is out of range gets NaN. */
int dida_letkf_obs_operator_device(dida_letkf_handle *h, const double *xb, const double *axb,
                                   int n_sites, const double *site_lat, const double *site_lon,
                                   const int *site_row_ptr, int n_rows, const int *row_var,
                                   const double *row_ph, const double *yo, double *yb,
                                   double *dep);

/* Fills the halo of a state on device arrays (xb layout, and the mean with
 * the axb layout; either may be NULL): periodic in longitude, the nearest
 * interior latitude row past either pole.  The reference gets its halo from
 * the model's halo exchange between cycles; device-resident cycling uses this
 * in its place.  Interior entries are not changed. */
int dida_letkf_fill_halo_device(dida_letkf_handle *h, double *x, double *ax);

/* Caller epilogue of run_da_parallel (module_da_parallel.F90:219-261) on
 * device arrays: polar mask, gamma1/gamma2 blend and spreads.  xa_out has the
 * xb layout (halo entries set to 0); spread_b/spread_a are (n, nlat, nlon).
 * grid_lat is taken from set_grid. */
int dida_letkf_epilogue_device(dida_letkf_handle *h, const double *xb, const double *axb,
                               const double *xa, const double *axa, double polar_lat,
                               double gamma1, double gamma2, double *xa_out, double *spread_b,
                               double *spread_a);

/* ------------------------------------------------------------------------
 * Serving: continuous batching of analysis jobs on one GPU.
 *
 * No counterpart in the reference (it runs one DA cycle per MPI job).  A job
 * is a configuration, a grid, a device-resident background and the obs of
 * n_cycles cycles; the server admits queued jobs in FIFO order while its HBM
 * budget and active-job limit allow, runs each admitted job on its own HIP
 * stream (jobs overlap on the device), keeps the ensemble on the device
 * between a job's cycles (the analysis of cycle c is the background of cycle
 * c+1), and records submit/start/end times.  A job's arithmetic does not
 * depend on what else runs: results are the same alone and in a mix.
 * ------------------------------------------------------------------------ */
typedef struct dida_serve dida_serve;

typedef struct dida_serve_config {
    size_t hbm_budget_bytes;  /* device memory admitted jobs may hold in total:
                                 allocated once by dida_serve_create (rounded
                                 down to 256 bytes); each admitted job gets a
                                 contiguous region of it (its size rounded up
                                 to 256 bytes, first fit, FIFO) from which
                                 every device buffer of the job is carved, so
                                 no device allocation or free runs while other
                                 jobs are in flight */
    int max_active;           /* jobs running at once (one stream each), >= 1 */
} dida_serve_config;

/* one cycle's obs set; pointers are device memory (as dida_letkf_set_obs_device).
 * Either yb and dep are given, or yb is NULL and row_var, row_ph and yo are:
 * then the server computes yb and dep from the cycle's background with
 * dida_letkf_obs_operator_device (the background of cycle c > 0 is the
 * analysis of cycle c - 1, still on the device). */
typedef struct dida_serve_obs {
    int n_sites;
    const double *site_lat, *site_lon;
    const int *site_row_ptr;
    int n_rows;
    const double *yb, *dep, *rinv, *vcoord;
    const int *row_var;
    const double *row_ph, *yo;
} dida_serve_obs;

typedef struct dida_serve_job {
    dida_letkf_config cfg;
    const float *grid_lat, *grid_lon;   /* host, copied at submit */
    const double *xb, *axb;             /* device, background of cycle 0 (read only) */
    int n_cycles;
    const dida_serve_obs *obs;          /* host array of n_cycles entries (copied at
                                           submit); the device arrays it points to must
                                           stay valid until the job has ended */
    double *xa, *axa;                   /* device, analysis of the last cycle */
} dida_serve_job;

enum {
    DIDA_JOB_QUEUED = 0,
    DIDA_JOB_RUNNING = 1,
    DIDA_JOB_DONE = 2,
    DIDA_JOB_FAILED = 3,
    DIDA_JOB_CANCELLED = 4
};

typedef struct dida_serve_status {
    int state;                /* DIDA_JOB_* */
    int rc;                   /* analysis return code (or error) of the job */
    double t_submit, t_start, t_end;   /* seconds on dida_serve_now()'s clock */
    long long n_analyses;     /* local analyses over all cycles */
    size_t bytes;             /* HBM the job was admitted with */
} dida_serve_status;

int dida_serve_create(const dida_serve_config *cfg, dida_serve **out);
/* Queues a job; DIDA_ERR_ADMISSION if it can never fit the budget. */
int dida_serve_submit(dida_serve *s, const dida_serve_job *job, long long *job_id);
int dida_serve_poll(dida_serve *s, long long job_id, dida_serve_status *st);
int dida_serve_wait(dida_serve *s, long long job_id, dida_serve_status *st);
/* Cancels a queued job; a running job stops after its current cycle. */
int dida_serve_cancel(dida_serve *s, long long job_id);
/* Cancels queued jobs, waits for running ones, frees the server. */
int dida_serve_destroy(dida_serve *s);
double dida_serve_now(void);
/* Device bytes one job of this shape needs (what admission charges). */
size_t dida_letkf_job_bytes(const dida_letkf_config *cfg, int n_sites, int n_rows);

const char *dida_letkf_last_error(void);
int dida_letkf_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DIDA_LETKF_H */
