// letkf_kernels.h -- kernel parameter blocks and launchers (host + device).
#pragma once
#include <hip/hip_runtime.h>

namespace dida {

constexpr int kMaxVars = 16;
// records the message returned by dida_letkf_last_error (per host thread)
int set_error(int code, const char* fmt, ...);
}  // namespace dida
struct dida_letkf_config;
namespace dida {
// analysis columns per batch whose scratch (with the row lists of n_rows
// obs rows, provisioned as rows_per_column_bound) fits in `bytes`
long long batch_columns_for_bytes(const dida_letkf_config& c, size_t bytes, int n_rows = 0);
long long rows_per_column_bound(const dida_letkf_config& c, int n_rows);
constexpr int kRotCapPerK2 = 2;   // rotation log: 2 k^2 Givens per analysis, plus
// the identity entries that align each chase to blocks of 8 positions
inline int rot_cap_of(int k) { return kRotCapPerK2 * k * k + 16 * k + 64; }
constexpr int kHdrCapPerK = 8;    // chase log: 8 k chases per analysis
// eigensolver status per analysis: 0 converged; l + 1 > 0: the QL of
// eigenvalue l did not converge; kStatusLogFull: the rotation or chase log
// filled up; kStatusNotRedone: flagged by the first pass when the list pass
// was full (the analysis keeps its background); kStatusNonFinite: overline_w
// is not finite (a NaN or Inf in d or R^-1 with a finite A; the analysis
// keeps its background)
constexpr int kStatusLogFull = -1;
constexpr int kStatusNotRedone = -2;
constexpr int kStatusNonFinite = -3;
// the update keeps the background of an analysis with this first-pass status
__host__ __device__ inline bool keeps_background(int status) {
  return status == kStatusNotRedone || status == kStatusNonFinite;
}

struct Geom {
  const float* glat;   // nlat, real(4) degrees
  const float* glon;   // nlon
  int nlat, nlon;
  int cube_n;
  int cube_s;          // cube start stride: cube_n - overlap
  int nlat_c, nlon_c;  // cubes per direction
};

// 1-based clipped bounds [s, e] of cube c along a direction of n points
__host__ __device__ inline void cube_span(const Geom& g, int c, int n, int* s, int* e) {
  *s = c * g.cube_s + 1;
  *e = *s + g.cube_n - 1 < n ? *s + g.cube_n - 1 : n;
}

struct SelectParams {
  Geom g;
  // sorted site index
  const double* s_cos;
  const double* s_sin;
  const double* s_lonr;
  const double* s_cosl;   // cos / sin of s_lonr
  const double* s_sinl;
  const int* s_row0;
  const int* s_nrow;
  const int* bin_start;   // nb_lat*nb_lon + 1
  double v_lo, v_hi;      // guard band of the support test in cos(angle)
  int nb_lat, nb_lon;
  double bdeg_lat, bdeg_lon;
  double theta_deg;       // support half-angle incl. margin, degrees
  double local_dist;
  int gc_filter;
  int h0, h1;             // horizontal analyses [h0, h1)
  long long* cnt;         // count pass: rows per column (batch-local)
  const long long* off;   // fill pass: exclusive offsets (batch-local)
  int* row_idx;
  double* row_gc;
  long long cap;          // fill pass: capacity of row_idx / row_gc
  // horizontal loop: the fill pass stores w = rinv * rho in row_gc and
  // w * dep in row_wd (weights = 1), or, for the BF16 products, the FP32
  // pair {sqrt(w), w * dep} in row_gc's 8 bytes and the row's offset in Yb,
  // row * k, in row_idx (weights = 2), so the products read them with the
  // list position
  int weights;
  int k;
  const float* rinv32;
  const double* dep;
  double* row_wd;
  // tile mode (k_select_tile): tiles of `tile` consecutive analyses of
  // [h0, h1); the union of their candidate rows, with the sorted site index
  int tile;
  int* row_site;
  // column mode: > 0 visits the columns in bands of band_w longitudes,
  // latitude by latitude (band_l latitudes per longitude)
  int band_w, band_l;
};

struct ProdParams {
  Geom g;
  int k, kp;
  int nzc, vertical, vcoord_mode, cube_nz, nz, halo, n, ps_index;
  double vert_dist;
  int gc_filter_vertical;
  const double* hyai;
  const double* hybi;
  double p0, ptop;
  const double* axb;
  const int* row_idx;        // obs rows (Yb offsets row * k for the BF16
                             // products when pre_weighted)
  const double* row_gc;      // GC weight, or w = rinv * rho when pre_weighted
                             // ({sqrt(w), w dep} in FP32 for the BF16 products)
  const double* row_wd;      // w * dep when pre_weighted
  int pre_weighted;
  const long long* row_off;  // per batch column, n_cols + 1 entries
  long long row_cap;         // entries of row_idx / row_gc / row_wd
  const double* yb64;
  const float* yb32;
  unsigned yb32_bytes;       // bytes of yb32's allocation (padding included),
                             // at most kYbOffsetLimit
  const double* dep;
  const float* rinv32;
  const double* vco;
  long long a0, n_an;        // analyses [a0, a0 + n_an)
  int h0;
  double* A;                 // batch-local, kp*kp per analysis
  double* b;                 // batch-local, kp per analysis
  const int* order;          // optional processing order (batch-local analysis offsets)
  int band_w, band_l;        // band_w > 0: visit bands of band_w longitudes latitude by
                             // latitude (band_l latitudes; batch of whole longitudes)
  int n_blocks;              // set by the launcher
  // tile mode (horizontal loop, k <= 64): row_off indexes tiles of `tile`
  // consecutive analyses; each union row carries its sorted site index
  int tile;
  const int* row_site;
  const double* s_cos;
  const double* s_sin;
  const double* s_lonr;
  double local_dist;
  int gc_filter;
  int* nobsl;                // tile mode: local obs rows per analysis (batch-local)
  // queue schedule: the launch's wavefront groups take analyses one at a
  // time from positions [seg_off[seg_lo], seg_off[seg_hi]) of `order` (the
  // length buckets, longest first), split in kQueueParts contiguous parts
  // with a counter each (q_counter[0 .. kQueueParts))
  const long long* seg_off;
  int seg_lo, seg_hi;
  int* q_counter;
  int groups;                // row groups of 32 loaded per step (1 or 2; k <= 64 BF16)
  // optional timing (queue schedule): wall-clock ticks summed per bucket
  // position, and per launch the earliest start, latest end and summed busy
  // ticks of its workers and the worker count (4 entries)
  unsigned long long* bucket_ticks;
  unsigned long long* launch_ticks;
};

// Length buckets over the local obs count p of a column: 0: p = 0,
// 1: 1..256, 2: 257..512, ... 6: 4097..8192, 7: more.  The order list holds
// the buckets longest first: bucket b at position kBuckets - 1 - b.
// Row lists and FP32 Yb are allocated kListPad entries (Yb: kListPad rows'
// worth of values) past their end, and the list tail holds valid row
// indices (0), so the BF16 products may read a group's positions past the end
// of an analysis' list (with weight 0) without clamping each index.
// The reduced-precision products read the FP32 Yb at 32-bit byte offsets
// below this limit; a position past the end of a row list reads at the limit
// itself, which lies past the buffer resource and returns 0.
constexpr unsigned kYbOffsetLimit = 0x7fff0000u;
constexpr int kListPad = 128;
constexpr int kBuckets = 8;
constexpr int kQueueParts = 8;   // queue counters per bucket position (one per XCD)
constexpr bool kQueueSplit = false;  // split a segment among the XCDs
__host__ __device__ inline int bucket_of(long long p) {
  if (p <= 0) return 0;
  if (p <= 256) return 1;
  int b = 1;
  long long e = 256;
  while (b < kBuckets - 1 && p > e) {
    e *= 2;
    ++b;
  }
  return b;
}

struct BucketParams {
  const int* counts;         // rows per column (batch-local), n_cols entries
  int n_cols, nzc;
  long long* blk;            // per block of kBucketBlock columns: kBuckets counts, then
                             // kBuckets row sums, then kBuckets offsets
  long long* seg_off;        // kBuckets + 1 positions (analyses) in the order list
  long long* bucket_stats;   // kBuckets analyses, kBuckets rows
  int* order;                // n_cols * nzc batch-local analysis offsets
  int* q_counters;           // zeroed by the scan kernel (kBuckets x kQueueParts)
  unsigned long long* ticks; // optional: kBuckets bucket sums, then 4 per bucket
                             // position for launch spans (reset by the scan kernel)
  int band_w, band_l;        // > 0: columns enter the sort in bands of band_w
                             // longitudes, latitude by latitude (band_index)
};
constexpr int kBucketBlock = 1024;

struct XformParams {
  Geom g;
  int k, kp;
  double inflation;
  int guard;
  int a_f32;                 // A in the BF16 products' FP32 tile order (a_entry)
  const double* A;
  const double* b;
  long long a0, n_an;
  int nzc, vertical, cube_nz, nz, halo, n, n_var;
  int var_start[kMaxVars];
  int var_lev[kMaxVars];
  const double* xb;          // NULL: transform only
  const double* axb;
  double* xa;
  double* axa;
  double* trans_out;         // optional, k*k per analysis (global index)
  double* w_out;             // optional, k per analysis
  int* status;               // optional, per analysis
  const int* list;           // optional: analyses to run (global indices)
  // FP32 state and fused epilogue, as in EigParams (split eigensolver only)
  int state_f32;
  const float* xb32;
  float* xa32;
  int epi;
  double polar_lat, gamma1, gamma2;
  double* xa_out;
  double* spread_b;
  double* spread_a;
};

struct EigParams {
  Geom g;
  int k, kp;
  double inflation;
  int guard;
  int max_iter;             // QL iterations per eigenvalue before the matrix is flagged
  long long a0, n_an;
  // Flag list.  First pass (list == NULL): k_ql_log appends every analysis it
  // flags to flag_list (global index, at most flag_cap entries; *flag_n
  // counts all of them, *flag_total accumulates over the batches of a call).
  // Rerun pass (list != NULL): the kernels take analysis list[slot] for
  // slot < min(*list_n, list_cap); batch-local inputs (d, e, refl, tau, b) are
  // indexed by list[slot] - a0, the pass's own scratch (rot, hdr, nchase, lam)
  // by slot.
  int* flag_list;
  int* flag_n;
  int* flag_total;
  int flag_cap;
  const int* list;
  const int* list_n;
  int list_cap;
  const double* A;          // batch-local, full symmetric kp x kp (a_entry)
  int a_f32;                // A in the BF16 products' FP32 tile order
  const double* b;          // batch-local, kp
  double* refl;             // batch-local, kp x kp reflectors (column j = v_j)
  double* tau;              // batch-local, kp
  double* d;                // batch-local, kp (tridiagonal diagonal)
  double* e;                // batch-local, kp (subdiagonal)
  double* lam;              // batch-local, kp (eigenvalues of A)
  double2* rot;             // batch-local, rot_cap Givens (c, s) per analysis
  int4* hdr;                // batch-local, hdr_cap chases (lo, hi, off) per analysis
  int rot_cap, hdr_cap;
  int tridiag_lds;          // k <= 64: tridiagonalise in LDS (1) or in VGPRs (0)
  int* nchase;              // batch-local
  int* status;              // global index
  int nzc, vertical, cube_nz, nz, halo, n, n_var;
  int var_start[kMaxVars];
  int var_lev[kMaxVars];
  const double* xb;
  const double* axb;
  double* xa;
  double* axa;
  double* trans_out;        // optional, global index
  double* w_out;            // optional, global index
  // FP32 state (state_f32 != 0): the perturbations are xb32/xa32 (member
  // fastest, as xb/xa); the means stay FP64 in axb/axa
  int state_f32;
  const float* xb32;
  float* xa32;
  // Fused caller epilogue (epi != 0, module_da_parallel.F90:219-261): instead
  // of xa/axa the update writes xa_out = full members
  // gamma1 (axb + xb) + gamma2 (axa + xa) (xb + axb where |lat| > polar_lat)
  // and spread_b/spread_a [lon][lat][n]
  int epi;
  double polar_lat, gamma1, gamma2;
  double* xa_out;
  double* spread_b;
  double* spread_a;
};

// observation operator + centring (gen_obs_struct)
struct ObsOpParams {
  int k, n, nz, nlat, nlon, halo, ps_index, n_var, n_rows;
  int var_start[kMaxVars];   // 1-based
  int var_lev[kMaxVars];
  const float* glat;         // nlat (closed by 90)
  const float* glon;         // nlon (closed by 360)
  const double* hyai;        // nz + 1
  const double* hybi;
  double p0, ptop;
  const double* xb;
  const double* axb;
  int n_sites;
  const double* site_lat;
  const double* site_lon;
  const int* site_row_ptr;
  const int* row_var;        // 0-based DA variable
  const double* row_ph;      // level pressure (Pa)
  const double* yo;
  double* yb;                // n_rows x k
  double* dep;
};

// The caller's split of the full ensemble state into mean and perturbations
// (module_da_parallel.F90:186-211): per state element, axb = sum_e x / ens and
// xb = x - axb (FP64, or rounded to FP32); elements with a NaN member are
// counted, and nan[1] keeps n_elems - (first such element).
struct SplitParams {
  const double* x_in;    // (k, n_elems), member fastest
  double* xb;            // FP64 perturbations, or null
  float* xb32;           // FP32 perturbations, or null
  double* axb;
  long long n_elems;
  int k;
  unsigned long long* nan;  // [count, n_elems - first]; zeroed by the launcher
};
hipError_t launch_split(hipStream_t st, const SplitParams& p);

// Overlapped cubes (letkf_compute.F90:211-255): each column's analysis is the
// mean of the analyses of the cubes that cover it.
struct OverlapParams {
  Geom g;
  int k, n, halo;
  const double* xb;
  const double* axb;
  double* xa;
  double* axa;
  const double* trans;   // per cube, Fortran (k, k): trans(m, e) at [c k^2 + e k + m]
  const int* status;     // per cube: nonzero keeps the background (trans = I)
};
hipError_t launch_overlap_update(hipStream_t st, const OverlapParams& p);

struct EpiParams {
  int k, n, nlat, nlon, halo;
  const float* glat;
  const double* xb;
  const double* axb;
  const double* xa;
  const double* axa;
  double polar_lat, gamma1, gamma2;
  double* xa_out;
  double* spread_b;
  double* spread_a;
};

hipError_t launch_site_keys(hipStream_t st, int n_sites, const double* lat, const double* lon,
                            double obs_polar, int nb_lat, int nb_lon, double bdeg_lat,
                            double bdeg_lon, int* key, int* val, int* hist);
hipError_t launch_site_gather(hipStream_t st, int n_sites, const int* order, const double* lat,
                              const double* lon, const int* row_ptr, double* s_cos, double* s_sin,
                              double* s_lonr, double* s_cosl, double* s_sinl, int* s_row0,
                              int* s_nrow);
hipError_t launch_select(hipStream_t st, const SelectParams& p, bool fill);
// exclusive prefix sum of n long longs (tmp: scan_tmp_elems(n) entries)
long long scan_tmp_elems(long long n);
hipError_t launch_exclusive_scan(hipStream_t st, const long long* in, long long* out, long long n,
                                 long long* tmp);
// after a batch's scan: per-list counts into counts[0..n_list) (off[i+1]-off[i]) and
// the batch's total off[n_list] into *total
hipError_t launch_list_counts(hipStream_t st, const long long* off, int n_list, int* counts,
                              long long* total);
hipError_t launch_zero_ll(hipStream_t st, long long* p, long long n);
hipError_t launch_select_tile(hipStream_t st, const SelectParams& p, bool fill);
hipError_t launch_products(hipStream_t st, const ProdParams& p, int precision);
// bucket schedule over the positions [p.seg_lo, p.seg_hi) of p.order: in
// dispatch order, or on a persistent work queue when p.q_counter is set
// (kQueueParts zeroed counters).  capacity: wavefront groups the device holds
// at once for this launch
hipError_t launch_products_queue(hipStream_t st, const ProdParams& p, int precision,
                                 long long* capacity = nullptr);
// length buckets of a batch's columns (BucketParams.blk: bucket_blk_elems)
long long bucket_blk_elems(int n_cols);
hipError_t launch_buckets(hipStream_t st, const BucketParams& p);
hipError_t launch_products_tile(hipStream_t st, const ProdParams& p, int precision);
hipError_t launch_transform(hipStream_t st, const XformParams& p);
hipError_t launch_epilogue(hipStream_t st, const EpiParams& p);
// ev: null, or 5 events recorded before the tridiagonalisation, the QL, the
// eigenvector/transform kernel and the update kernel, and after
hipError_t launch_obs_operator(hipStream_t st, const ObsOpParams& p);
// halo of a [lon + 2h][lat + 2h][n] field with `inner` values per element
hipError_t launch_fill_halo(hipStream_t st, double* x, int nlat, int nlon, int halo, long long inner);
// zeroes the halo columns of such a field
hipError_t launch_zero_halo(hipStream_t st, double* x, int nlat, int nlon, int halo, long long inner);
hipError_t launch_eig(hipStream_t st, const EigParams& p, hipEvent_t* ev = nullptr);
// second pass over the analyses the first flagged (p.list / p.list_n / p.list_cap)
hipError_t launch_eig_rerun(hipStream_t st, const EigParams& p);
hipError_t launch_zero_ints(hipStream_t st, int* p, int n);
// k = 65..128 batches of at least 2 kEigPipeMinHalf analyses: the batch in two
// halves, the second on st2 (ordered after the first half's
// tridiagonalisation, joined back into st through e_join); ev as above, with
// the first half's kernel boundaries (the second half's kernels run beside
// them).  Smaller batches and k <= 64 take launch_eig on st.
constexpr long long kEigPipeMinHalf = 1024;
hipError_t launch_eig_pipelined(hipStream_t st, hipStream_t st2, hipEvent_t e_first,
                                hipEvent_t e_join, const EigParams& p, hipEvent_t* ev = nullptr);

// Entry (r, c) of one analysis' A, whose slot holds kp * kp doubles.  FP64:
// column-major.  a_f32 (the BF16 products, kp = 32 or 64): the FP32 sums in
// the products' tile order, column-major in the first half of the slot,
// where member m = T i + t (T = kp / 16, i < 16) sits at position 16 t + i.
__device__ __forceinline__ int a_tile_pos(int m, int kp) {
  const int sh = kp == 64 ? 2 : 1;
  return ((m & ((1 << sh) - 1)) << 4) + (m >> sh);
}
__device__ __forceinline__ double a_entry(const double* Ag, int a_f32, int kp, int r, int c) {
  if (a_f32)
    return (double)reinterpret_cast<const float*>(Ag)[a_tile_pos(c, kp) * kp + a_tile_pos(r, kp)];
  return Ag[(long long)c * kp + r];
}

}  // namespace dida
