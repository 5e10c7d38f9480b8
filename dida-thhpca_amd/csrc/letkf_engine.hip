// letkf_engine.hip -- host runtime behind the C ABI (include/dida_letkf.h).
//
// Owns device state per handle: grid, the binned obs-site index, obs rows in
// the storage precision of the configured mode, and per-batch work buffers.
// One analysis call runs, per batch of horizontal columns:
//   select(count) -> exclusive scan -> select(fill) -> products -> transform
// on the handle's stream.  No host synchronisation happens inside a batch
// except to read the batch's total row count (to size the row list).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <unistd.h>

#include "dida_letkf.h"
#include "letkf_device.h"
#include "letkf_kernels.h"
#include "letkf_pool.h"

using namespace dida;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace dida {
int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
}  // namespace dida

namespace {

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return set_error(e_ == hipErrorOutOfMemory ? DIDA_ERR_NOMEM : DIDA_ERR_HIP,          \
                       "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,    \
                       __LINE__);                                                          \
  } while (0)

#define RC_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ < 0) return rc_;    \
  } while (0)

thread_local DevicePool* tl_pool = nullptr;

// Test hook: with DIDA_FAIL_ALLOC_AFTER=N set, the device allocations of a
// host thread after its first N since the value was set fail (NOMEM).
bool alloc_refused() {
  thread_local std::string seen;
  thread_local long long count = 0;
  const char* v = std::getenv("DIDA_FAIL_ALLOC_AFTER");
  if (!v || !*v) {
    seen.clear();
    return false;
  }
  if (seen != v) {
    seen = v;
    count = 0;
  }
  return count++ >= std::atoll(v);
}

// Growable device buffer, or (wrap) a fixed piece of memory it does not own:
// ensure() then succeeds only within the piece and never allocates.  While a
// DevicePool is bound to the thread, growth takes blocks of that pool
// instead of hipMalloc (and gives them back instead of hipFree).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DevicePool* pool = nullptr;     // the pool p came from, if any
  int ensure(size_t n) {
    if (n <= bytes && p) return DIDA_OK;
    if (!owned)
      return set_error(DIDA_ERR_ARG, "fixed scratch piece of %zu bytes is too small (%zu needed)",
                       bytes, n);
    release();
    if (n == 0) n = 16;
    if (alloc_refused())
      return set_error(DIDA_ERR_NOMEM, "device allocation of %zu bytes refused "
                       "(DIDA_FAIL_ALLOC_AFTER)", n);
    if (DevicePool* pl = tl_pool) {
      const size_t off = pl->fit.alloc(n);
      if (off == FitAlloc::npos)
        return set_error(DIDA_ERR_NOMEM, "job memory region exhausted: %zu bytes needed, "
                         "largest free block %zu of %zu", n, pl->fit.largest(), pl->fit.size());
      p = pl->base + off;
      pool = pl;
      bytes = n;
      return DIDA_OK;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return set_error(DIDA_ERR_NOMEM, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return DIDA_OK;
  }
  void wrap(void* q, size_t n) {
    release();
    p = q;
    bytes = n;
    owned = false;
  }
  void release() {
    if (p && owned) {
      if (pool) (void)pool->fit.release((size_t)((char*)p - pool->base), bytes);
      else (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    owned = true;
    pool = nullptr;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace

namespace dida {
DevicePool* bind_thread_pool(DevicePool* pool) {
  DevicePool* old = tl_pool;
  tl_pool = pool;
  return old;
}
DevicePool* thread_pool() { return tl_pool; }
}  // namespace dida

namespace {

__global__ void k_f64_to_f32(long long n, const double* __restrict__ in, float* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// out[i] = i for i < n, then 0 for the list padding up to n + pad
__global__ void k_iota(long long n, long long pad, int* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n + pad) out[i] = i < n ? (int)i : 0;
}

int kp_of(int k) { return k <= 32 ? 32 : (k <= 64 ? 64 : 128); }

// Layout of the end-of-call summary, in 8-byte words: the flagged analyses,
// the local obs rows (all levels), the largest column count, the columns over
// 10,000 rows, the two eigensolver flag counters, the two split NaN counters,
// then the per-batch row totals, bucket statistics and bucket ticks.
constexpr int kSumHead = 8;

struct SummaryParams {
  const int* status;
  long long n_an;
  const int* counts;
  long long n_h;
  int nzc;
  const int* flags;
  const unsigned long long* nan;      // null without the device split
  const long long* tot;
  int n_tot;
  const long long* bst;
  int n_bst;
  const unsigned long long* tk;
  int n_tk;
  long long* out;
};

// One workgroup of 1024 threads gathers a call's results into one small
// array, so that the host reads them back with one copy.
__global__ __launch_bounds__(1024) void k_call_summary(SummaryParams p) {
  __shared__ long long red[3][16];
  __shared__ int red_max[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  long long flagged = 0, rows = 0, over = 0;
  int mx = 0;
  for (long long i = t; i < p.n_an; i += 1024) flagged += p.status[i] != 0;
  for (long long i = t; i < p.n_h; i += 1024) {
    const int r = p.counts[i];
    rows += (long long)r * p.nzc;
    mx = max(mx, r);
    over += r > 10000;
  }
  for (int o = 32; o > 0; o >>= 1) {
    flagged += __shfl_xor(flagged, o, 64);
    rows += __shfl_xor(rows, o, 64);
    over += __shfl_xor(over, o, 64);
    mx = max(mx, __shfl_xor(mx, o, 64));
  }
  if (lane == 0) {
    red[0][w] = flagged;
    red[1][w] = rows;
    red[2][w] = over;
    red_max[w] = mx;
  }
  __syncthreads();
  if (t == 0) {
    long long a = 0, b = 0, c = 0;
    int m = 0;
    for (int i = 0; i < 16; ++i) {
      a += red[0][i];
      b += red[1][i];
      c += red[2][i];
      m = max(m, red_max[i]);
    }
    p.out[0] = a;
    p.out[1] = b;
    p.out[2] = m;
    p.out[3] = c;
    p.out[4] = p.flags[0];
    p.out[5] = p.flags[1];
    p.out[6] = p.nan ? (long long)p.nan[0] : 0;
    p.out[7] = p.nan ? (long long)p.nan[1] : 0;
  }
  long long* o = p.out + kSumHead;
  for (int i = t; i < p.n_tot; i += 1024) o[i] = p.tot[i];
  o += p.n_tot;
  for (int i = t; i < p.n_bst; i += 1024) o[i] = p.bst[i];
  o += p.n_bst;
  for (int i = t; i < p.n_tk; i += 1024) o[i] = (long long)p.tk[i];
}

}  // namespace

// Scratch of the split eigensolver for one batch of analyses.
struct EigScratch {
  DevBuf refl, tau, d, e, lam, rot, hdr, nchase;
  // flags: [0] analyses the batch's first pass flagged, [1] the same summed
  // over the batches of a call, [2 ..] their list; rr_*: the list pass's own
  // scratch (its longer rotation log, chase headers, eigenvalues)
  DevBuf flags, rr_rot, rr_hdr, rr_nchase, rr_lam;
  std::vector<int> st_host;
  // second stream and events of the k = 65..128 half-batch pipeline, created
  // on first use (default priority)
  hipStream_t st2 = nullptr;
  hipEvent_t e_first = nullptr, e_join = nullptr;
  void release_streams() {
    if (st2) {
      (void)hipStreamSynchronize(st2);
      (void)hipStreamDestroy(st2);
      st2 = nullptr;
    }
    if (e_first) (void)hipEventDestroy(e_first);
    if (e_join) (void)hipEventDestroy(e_join);
    e_first = e_join = nullptr;
  }
};


// Tridiagonalisation variant: DIDA_TRIDIAG=lds (1) selects the LDS-resident
// kernel, anything else the register-resident one (faster at every k).
int tridiag_variant() {
  static const int v = [] {
    const char* e = std::getenv("DIDA_TRIDIAG");
    return (e && std::strcmp(e, "lds") == 0) ? 1 : 0;
  }();
  return v;
}

// Analyses per tile of the tile-union products (config products_kernel = 2,
// horizontal loop, k <= 64).
constexpr int kTile = 8;

// Visiting order of the products kernels: bands of w longitudes latitude by
// latitude (w = 32 unless DIDA_BAND sets it; 0: one meridian after another).
int products_band() {
  static const int v = [] {
    const char* e = std::getenv("DIDA_BAND");
    return e ? std::max(0, std::atoi(e)) : 32;
  }();
  return v;
}


// bytes of split-eigensolver scratch per analysis
size_t eig_bytes_per_analysis(int k) {
  const int kp = kp_of(k);
  return sizeof(double) * ((size_t)kp * kp + 4 * kp) + 16 * (size_t)rot_cap_of(k) +
         16 * (size_t)kHdrCapPerK * k + sizeof(int);
}

// QL iterations per eigenvalue before a matrix is flagged (first pass; the
// list pass allows twice as many), as the EISPACK/Numerical Recipes QL does
constexpr int kQlMaxIter = 30;

// Analyses the list pass can redo per batch, and its per-analysis scratch
// (a rotation log and chase log twice the first pass's)
long long rerun_cap_of(int k, long long n_an) {
  return std::min<long long>(n_an, k > 64 ? 128 : 256);
}
size_t rerun_bytes_per_slot(int k) {
  const int kp = kp_of(k);
  return 16 * (size_t)(2 * rot_cap_of(k)) + 16 * (size_t)(2 * kHdrCapPerK * k) + sizeof(int) +
         sizeof(double) * kp + sizeof(int);
}

// Optional positive integer from the environment (0 when unset).  The
// eigensolver's limits can be lowered this way to make the flag paths run:
// DIDA_ROT_CAP / DIDA_QL_MAXITER (first pass), DIDA_RERUN_ROT_CAP /
// DIDA_RERUN_MAXITER (list pass).
int env_int(const char* name) {
  const char* e = std::getenv(name);
  if (!e || !*e) return 0;
  return std::max(0, std::atoi(e));
}

// Sizes every eigensolver buffer for batches of up to n analyses (the list
// pass's too), so that the launches of a call allocate nothing.
int eig_prepare(EigScratch& s, int k, long long n) {
  const int kp = kp_of(k);
  const long long cap = rerun_cap_of(k, n);
  RC_TRY(s.refl.ensure(sizeof(double) * n * kp * kp));
  RC_TRY(s.tau.ensure(sizeof(double) * n * kp));
  RC_TRY(s.d.ensure(sizeof(double) * n * kp));
  RC_TRY(s.e.ensure(sizeof(double) * n * kp));
  RC_TRY(s.lam.ensure(sizeof(double) * n * kp));
  RC_TRY(s.rot.ensure((size_t)16 * n * rot_cap_of(k)));
  RC_TRY(s.hdr.ensure((size_t)16 * n * kHdrCapPerK * k));
  RC_TRY(s.nchase.ensure(sizeof(int) * n));
  RC_TRY(s.flags.ensure(sizeof(int) * (2 + cap)));
  RC_TRY(s.rr_rot.ensure((size_t)16 * cap * 2 * rot_cap_of(k)));
  RC_TRY(s.rr_hdr.ensure((size_t)16 * cap * 2 * kHdrCapPerK * k));
  RC_TRY(s.rr_nchase.ensure(sizeof(int) * cap));
  RC_TRY(s.rr_lam.ensure(sizeof(double) * cap * kp));
  return DIDA_OK;
}

// Zeroes the call-wide flag count; run once per analysis call.
int eig_begin_call(hipStream_t st, EigScratch& s, int k, long long max_batch) {
  RC_TRY(s.flags.ensure(sizeof(int) * (2 + rerun_cap_of(k, max_batch))));
  HIP_TRY(launch_zero_ints(st, s.flags.as<int>(), 2));
  return DIDA_OK;
}

// Eigensolver parameters of the batch xp (its scratch allocated): the split
// eigensolver (k_tridiag, k_ql_log, k_vec + k_update) over analyses
// [xp.a0, xp.a0 + xp.n_an).
int prepare_eig(EigScratch& s, const XformParams& xp, EigParams& ep) {
  const long long n = xp.n_an;
  const int k = xp.k, kp = xp.kp;
  int rot_cap = rot_cap_of(k);
  if (int v = env_int("DIDA_ROT_CAP")) rot_cap = std::max(16, std::min(v, rot_cap));
  const int hdr_cap = kHdrCapPerK * k;
  RC_TRY(s.refl.ensure(sizeof(double) * n * kp * kp));
  RC_TRY(s.tau.ensure(sizeof(double) * n * kp));
  RC_TRY(s.d.ensure(sizeof(double) * n * kp));
  RC_TRY(s.e.ensure(sizeof(double) * n * kp));
  RC_TRY(s.lam.ensure(sizeof(double) * n * kp));
  RC_TRY(s.rot.ensure((size_t)16 * n * rot_cap));
  RC_TRY(s.hdr.ensure((size_t)16 * n * hdr_cap));
  RC_TRY(s.nchase.ensure(sizeof(int) * n));
  const long long rcap = rerun_cap_of(k, n);
  if (s.flags.bytes < sizeof(int) * (2 + rcap))
    return set_error(DIDA_ERR_STATE, "eigensolver flags not set up for this batch");
  ep = EigParams{};
  ep.g = xp.g;
  ep.k = k;
  ep.kp = kp;
  ep.inflation = xp.inflation;
  ep.guard = xp.guard;
  ep.max_iter = env_int("DIDA_QL_MAXITER") ? env_int("DIDA_QL_MAXITER") : kQlMaxIter;
  ep.flag_n = s.flags.as<int>();
  ep.flag_total = s.flags.as<int>() + 1;
  ep.flag_list = s.flags.as<int>() + 2;
  ep.flag_cap = (int)rcap;
  ep.list = nullptr;
  ep.list_n = nullptr;
  ep.list_cap = 0;
  ep.a0 = xp.a0;
  ep.n_an = n;
  ep.A = xp.A;
  ep.a_f32 = xp.a_f32;
  ep.b = xp.b;
  ep.refl = s.refl.as<double>();
  ep.tau = s.tau.as<double>();
  ep.d = s.d.as<double>();
  ep.e = s.e.as<double>();
  ep.lam = s.lam.as<double>();
  ep.rot = s.rot.as<double2>();
  ep.hdr = s.hdr.as<int4>();
  ep.rot_cap = rot_cap;
  ep.hdr_cap = hdr_cap;
  // rows (k <= 64) or half rows (k_tridiag128) of A in registers unless
  // DIDA_TRIDIAG=lds asks for the LDS-resident kernel
  ep.tridiag_lds = tridiag_variant() == 1;
  ep.nchase = s.nchase.as<int>();
  ep.status = xp.status;
  ep.nzc = xp.nzc;
  ep.vertical = xp.vertical;
  ep.cube_nz = xp.cube_nz;
  ep.nz = xp.nz;
  ep.halo = xp.halo;
  ep.n = xp.n;
  ep.n_var = xp.n_var;
  for (int v = 0; v < kMaxVars; ++v) {
    ep.var_start[v] = xp.var_start[v];
    ep.var_lev[v] = xp.var_lev[v];
  }
  ep.xb = xp.xb;
  ep.axb = xp.axb;
  ep.xa = xp.xa;
  ep.axa = xp.axa;
  ep.trans_out = xp.trans_out;
  ep.w_out = xp.w_out;
  ep.state_f32 = xp.state_f32;
  ep.xb32 = xp.xb32;
  ep.xa32 = xp.xa32;
  ep.epi = xp.epi;
  ep.polar_lat = xp.polar_lat;
  ep.gamma1 = xp.gamma1;
  ep.gamma2 = xp.gamma2;
  ep.xa_out = xp.xa_out;
  ep.spread_b = xp.spread_b;
  ep.spread_a = xp.spread_a;
  return DIDA_OK;
}

// The list pass over the analyses the first pass flagged: k_ql_log with a
// rotation log and an iteration limit twice the first pass's, then k_vec and
// k_update, all on st after the first pass (no host synchronisation).
int rerun_flagged(hipStream_t st, EigScratch& s, const EigParams& ep) {
  const int k = ep.k, kp = ep.kp;
  const long long cap = ep.flag_cap;
  if (cap <= 0) return DIDA_OK;
  int rot_cap = 2 * rot_cap_of(k);
  if (int v = env_int("DIDA_RERUN_ROT_CAP")) rot_cap = std::max(16, std::min(v, rot_cap));
  const int hdr_cap = 2 * kHdrCapPerK * k;
  RC_TRY(s.rr_rot.ensure((size_t)16 * cap * rot_cap));
  RC_TRY(s.rr_hdr.ensure((size_t)16 * cap * hdr_cap));
  RC_TRY(s.rr_nchase.ensure(sizeof(int) * cap));
  RC_TRY(s.rr_lam.ensure(sizeof(double) * cap * kp));
  EigParams rp = ep;
  rp.list = ep.flag_list;
  rp.list_n = ep.flag_n;
  rp.list_cap = (int)cap;
  rp.max_iter = env_int("DIDA_RERUN_MAXITER") ? env_int("DIDA_RERUN_MAXITER") : 2 * kQlMaxIter;
  rp.rot = s.rr_rot.as<double2>();
  rp.rot_cap = rot_cap;
  rp.hdr = s.rr_hdr.as<int4>();
  rp.hdr_cap = hdr_cap;
  rp.nchase = s.rr_nchase.as<int>();
  rp.lam = s.rr_lam.as<double>();
  HIP_TRY(launch_eig_rerun(st, rp));
  return DIDA_OK;
}

// The half-batch pipeline's second stream and its fork / join events (a
// captured call records them as graph dependencies).
int eig_second_stream(EigScratch& s) {
  if (s.st2) return DIDA_OK;
  HIP_TRY(hipStreamCreateWithFlags(&s.st2, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&s.e_first, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&s.e_join, hipEventDisableTiming));
  return DIDA_OK;
}

// Transform + update of analyses [xp.a0, xp.a0 + xp.n_an): the split
// eigensolver, then its list pass over the analyses it flagged; mode 1 runs
// only the single kernel (k <= 64).  Mode 0 runs batches of 2 kEigPipeMinHalf
// analyses or more as two half-batches on two streams; mode 2 keeps the split
// eigensolver on st (each kernel timed by itself).
// kev: null, or 5 events for the split kernels (recorded only if *split is set)
int run_transform(hipStream_t st, EigScratch& s, const XformParams& xp, int mode,
                  hipEvent_t* kev = nullptr, bool* split = nullptr) {
  if (split) *split = false;
  // the single-kernel transform updates an FP64 state only
  if (mode == 1 && xp.k <= 64 && !xp.state_f32 && !xp.epi) {
    HIP_TRY(launch_transform(st, xp));
    return DIDA_OK;
  }
  EigParams ep;
  RC_TRY(prepare_eig(s, xp, ep));
  HIP_TRY(launch_zero_ints(st, ep.flag_n, 1));
  if (mode == 0 && xp.n_an >= 2 * kEigPipeMinHalf) {
    RC_TRY(eig_second_stream(s));
    HIP_TRY(launch_eig_pipelined(st, s.st2, s.e_first, s.e_join, ep, kev));
  } else {
    HIP_TRY(launch_eig(st, ep, kev));
  }
  if (split) *split = true;
  return rerun_flagged(st, s, ep);
}

// cubes along n grid points with starts every cube_n - overlap (cube_n
// without overlap): letkf_compute.F90:53,61 and :197,205
static long long cubes_along(int n, int cube_n, int overlap) {
  const int s = overlap > 0 ? cube_n - overlap : cube_n;
  return n >= 1 && s >= 1 ? (n - 1) / s + 1 : 0;
}
static long long cubes_h(const dida_letkf_config& c) {
  return cubes_along(c.nlat, c.cube_n, c.overlap) * cubes_along(c.nlon, c.cube_n, c.overlap);
}

namespace dida {
// columns whose per-batch scratch fits in `bytes`
long long batch_columns_for_bytes(const dida_letkf_config& c, size_t bytes, int n_rows) {
  const int kp = kp_of(c.k);
  const int cz = c.cube_nz > 0 ? c.cube_nz : c.nz;
  const long long nzc = c.vert_dist > 0.0 ? (c.nz + cz - 1) / cz : 1;
  const long long per_col =
      nzc * ((long long)kp * (kp + 1) * 8 + (long long)eig_bytes_per_analysis(c.k)) +
      20 * rows_per_column_bound(c, n_rows);
  return std::max<long long>(1, (long long)(bytes / (size_t)per_col));
}

// Local obs rows of a column to provision for when the obs count is known
// beforehand (serving): six times the rows of uniformly spread obs within the
// support radius 2c, plus 64; 0 for n_rows = 0.
long long rows_per_column_bound(const dida_letkf_config& c, int n_rows) {
  if (n_rows <= 0) return 0;
  const double th = std::min(M_PI, 2.0 * c.local_dist / kEarthF32);
  const double frac = std::min(1.0, 6.0 * 0.5 * (1.0 - std::cos(th)));
  return std::min<long long>(n_rows, (long long)(frac * n_rows) + 64);
}
}  // namespace dida

// Columns per batch: the per-batch scratch (products + eigensolver) is
// bounded to 16 GiB (a 360x180 k=64 cycle is one batch), or by max_batch_columns.
long long max_batch_columns(const dida_letkf_config& c) {
  long long mb = batch_columns_for_bytes(c, 16ull << 30);
  if (c.max_batch_columns > 0) mb = std::min<long long>(mb, c.max_batch_columns);
  const long long nh = cubes_h(c);
  return std::min<long long>(mb, nh);
}

struct dida_letkf_handle {
  dida_letkf_config cfg;
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int n = 0, nlat_c = 0, nlon_c = 0, nzc = 1, kp = 32;
  long long n_h = 0, n_an = 0;
  std::vector<double> hyai, hybi;
  bool have_grid = false, have_obs = false;
  // grid
  DevBuf glat, glon;
  // raw obs staging (device)
  DevBuf o_lat, o_lon, o_rowptr, o_yb64, o_dep, o_rinv64, o_vco;
  // obs index and rows
  int n_sites = 0, n_rows = 0;
  int nb_lat = 0, nb_lon = 0;
  double bdeg_lat = 0, bdeg_lon = 0, theta_deg = 0;
  DevBuf key_in, key_out, val_in, val_out, hist, bin_start, sort_tmp;
  DevBuf s_cos, s_sin, s_lonr, s_cosl, s_sinl, s_row0, s_nrow;
  DevBuf yb32, rinv32;
  DevBuf d_hyai, d_hybi;
  DevBuf row_site, nobsl_dev;         // tile-mode selection and row counts
  // work
  DevBuf cnt, off, scan_tmp, row_idx, row_gc, row_wd, Abuf, bbuf, status, trans, wout;
  DevBuf x_xb, x_axb, x_xa, x_axa;   // host-entry staging
  EigScratch eig;
  std::vector<int> nobsl_host;
  long long max_batch = 0;
  bool tiled = false;                 // tile-union products
  // row lists: capacity per batch (grown when a call finds more rows, which
  // then reruns), per-batch totals and per-column counts read after a call
  long long row_cap = 0;
  DevBuf batch_tot;
  DevBuf call_sum;                    // end-of-call summary (k_call_summary)
  // per-batch timing events (9 per batch), created once
  std::vector<hipEvent_t> ev;
  // hipGraph of a call's launches, replayed while every pointer and size it
  // captured is unchanged
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uintptr_t> gkey;
  hipEvent_t g_ev[2] = {nullptr, nullptr};
  // options (dida_letkf_set_option)
  int opt_schedule = 1, opt_timing = 0, opt_tune = 0, opt_graph = 0;
  // length buckets: order list, block counts, segment starts, queue counters,
  // per-batch bucket statistics and timing ticks
  DevBuf b_order, b_blk, b_seg, b_q, b_stats, b_ticks;
  DevBuf sp_xb, sp_xb32, sp_axb, sp_nan;  // device split of the full state
  int variant[kBuckets] = {0};        // products variant per bucket position
  std::vector<bool> split_timed;      // per batch: split eigensolver events recorded
  std::vector<long long> run_capacity;  // per batch and bucket position: the wavefront
                                        // groups the device held for the launch starting there
  std::string arch;                   // device architecture name (tuning cache key)
  double clock_khz = 100000.0;        // device wall clock (in-kernel timing)
};

// products variant per bucket position, from the tuning cache
static void resolve_variants(dida_letkf_handle* h);

extern "C" {

const char* dida_letkf_last_error(void) { return g_last_error.c_str(); }
int dida_letkf_abi_version(void) { return DIDA_LETKF_ABI_VERSION; }

void dida_letkf_config_default(dida_letkf_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->k = 20;
  c->nlat = 1;
  c->nlon = 1;
  c->nz = 1;
  c->n2d = 1;
  c->n3d = 0;
  c->halo = 0;
  c->cube_n = 1;
  c->cube_nz = 32;
  c->local_dist = 0.0;
  c->vert_dist = 0.0;
  c->inflation_factor = 1.0;
  c->obs_polar = 91.0;
  c->gc_filter = 0;
  c->gc_filter_vertical = 0;
  c->precision = DIDA_PREC_FP64;
  c->vcoord_mode = DIDA_VCOORD_LEVEL;
  c->ps_index = 0;
  c->p0 = 1.0e5;
  c->ptop = 0.0;
}

long long dida_letkf_n_analyses(const dida_letkf_config* c) {
  if (!c || c->cube_n < 1 || c->nlat < 1 || c->nlon < 1) return -1;
  if (c->overlap < 0 || c->overlap >= c->cube_n) return -1;
  long long nh = cubes_h(*c);
  if (c->vert_dist > 0.0) {
    int cz = c->cube_nz > 0 ? c->cube_nz : c->nz;
    nh *= (c->nz + cz - 1) / cz;
  }
  return nh;
}

int dida_letkf_create(const dida_letkf_config* cfg, dida_letkf_handle** out) {
  if (!cfg || !out) return set_error(DIDA_ERR_ARG, "null argument");
  *out = nullptr;
  const dida_letkf_config& c = *cfg;
  if (c.k < 2 || c.k > 128) return set_error(DIDA_ERR_ARG, "k=%d outside 2..128", c.k);
  if (c.nlat < 1 || c.nlon < 1 || c.nz < 1 || c.n2d < 0 || c.n3d < 0 || c.n2d + c.n3d < 1)
    return set_error(DIDA_ERR_ARG, "bad grid/variable sizes");
  if (c.n2d + c.n3d > kMaxVars) return set_error(DIDA_ERR_ARG, "more than %d variables", kMaxVars);
  if (c.halo < 0 || c.cube_n < 1) return set_error(DIDA_ERR_ARG, "bad halo/cube_n");
  if (c.overlap < 0 || c.overlap >= c.cube_n)
    return set_error(DIDA_ERR_ARG, "overlap=%d outside 0..cube_n-1", c.overlap);
  if (c.overlap > 0 && c.vert_dist > 0.0)
    return set_error(DIDA_ERR_UNSUPPORTED,
                     "overlapped cubes run the horizontal loop only (letkf.F90:227-235)");
  if (!(c.local_dist > 0.0)) return set_error(DIDA_ERR_ARG, "local_dist must be > 0");
  if (!(c.inflation_factor > 0.0)) return set_error(DIDA_ERR_ARG, "inflation_factor must be > 0");
  if (c.precision < DIDA_PREC_FP64 || c.precision > DIDA_PREC_BF16X3)
    return set_error(DIDA_ERR_ARG, "bad precision %d", c.precision);
  if (c.products_kernel < 0 || c.products_kernel > 2)
    return set_error(DIDA_ERR_ARG, "bad products_kernel %d", c.products_kernel);
  if (c.vert_dist > 0.0) {
    if (c.vcoord_mode == DIDA_VCOORD_LNP && (!c.hyai || !c.hybi))
      return set_error(DIDA_ERR_ARG, "LNP vertical mode needs hyai/hybi");
    if (c.vcoord_mode != DIDA_VCOORD_LEVEL && c.vcoord_mode != DIDA_VCOORD_LNP)
      return set_error(DIDA_ERR_ARG, "bad vcoord_mode");
    if (c.vcoord_mode == DIDA_VCOORD_LNP && (c.ps_index < 0 || c.ps_index >= c.n2d + c.n3d * c.nz))
      return set_error(DIDA_ERR_ARG, "bad ps_index");
  }
  dida_letkf_handle* h = new (std::nothrow) dida_letkf_handle();
  if (!h) return set_error(DIDA_ERR_NOMEM, "host allocation failed");
  h->cfg = c;
  h->cfg.hyai = nullptr;
  h->cfg.hybi = nullptr;
  if (c.hyai && c.hybi) {
    h->hyai.assign(c.hyai, c.hyai + c.nz + 1);
    h->hybi.assign(c.hybi, c.hybi + c.nz + 1);
  }
  if (h->cfg.cube_nz < 1) h->cfg.cube_nz = c.nz;
  h->n = c.n2d + c.n3d * c.nz;
  h->nlat_c = (int)cubes_along(c.nlat, c.cube_n, c.overlap);
  h->nlon_c = (int)cubes_along(c.nlon, c.cube_n, c.overlap);
  h->nzc = c.vert_dist > 0.0 ? (c.nz + h->cfg.cube_nz - 1) / h->cfg.cube_nz : 1;
  h->n_h = (long long)h->nlat_c * h->nlon_c;
  h->n_an = h->n_h * h->nzc;
  h->kp = kp_of(c.k);
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return set_error(DIDA_ERR_HIP, "HIP init failed: %s", hipGetErrorString(e));
  }
  h->stream = h->own_stream;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) {
      h->arch = prop.gcnArchName;
      const size_t cut = h->arch.find(':');
      if (cut != std::string::npos) h->arch.resize(cut);
    }
    if (h->arch.empty()) h->arch = "unknown";
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device) == hipSuccess &&
        khz > 0)
      h->clock_khz = khz;
  }
  // obs-space support angle 2c/a (radians -> degrees) with margin for the
  // real(4) pi scaling of the reference's distance
  const double a = 6371.39306640625;
  h->theta_deg = (2.0 * c.local_dist / a) * (180.0 / M_PI) * (1.0 + 1e-6) + 1e-2;
  double bd = std::min(10.0, std::max(0.5, h->theta_deg / 2.0));
  h->nb_lat = (int)std::ceil(180.0 / bd);
  h->nb_lon = (int)std::ceil(360.0 / bd);
  h->bdeg_lat = 180.0 / h->nb_lat;
  h->bdeg_lon = 360.0 / h->nb_lon;
  h->max_batch = max_batch_columns(c);
  h->tiled = h->nzc == 1 && h->kp <= 64 && c.products_kernel == 2;
  if (h->tiled) {
    // tiles never straddle batches, so results do not depend on the batching
    h->max_batch = std::max<long long>(kTile, (h->max_batch / kTile) * kTile);
  }
  resolve_variants(h);
  *out = h;
  return DIDA_OK;
}

int dida_letkf_destroy(dida_letkf_handle* h) {
  if (!h) return DIDA_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->eig.release_streams();
  DevBuf* bufs[] = {&h->glat, &h->glon, &h->o_lat, &h->o_lon, &h->o_rowptr, &h->o_yb64,
                    &h->o_dep, &h->o_rinv64, &h->o_vco, &h->key_in, &h->key_out, &h->val_in,
                    &h->val_out, &h->hist, &h->bin_start, &h->sort_tmp, &h->s_cos, &h->s_sin,
                    &h->s_lonr, &h->s_cosl, &h->s_sinl, &h->s_row0, &h->s_nrow, &h->yb32,
                    &h->rinv32, &h->d_hyai,
                    &h->d_hybi, &h->cnt, &h->off, &h->scan_tmp, &h->row_idx, &h->row_gc,
                    &h->row_site, &h->nobsl_dev, &h->row_wd,
                    &h->Abuf, &h->bbuf, &h->status, &h->trans, &h->wout, &h->x_xb, &h->x_axb,
                    &h->x_xa, &h->x_axa, &h->eig.refl, &h->eig.tau, &h->eig.d, &h->eig.e,
                    &h->eig.lam, &h->eig.rot, &h->eig.hdr, &h->eig.nchase,
                    &h->eig.flags, &h->eig.rr_rot, &h->eig.rr_hdr, &h->eig.rr_nchase,
                    &h->eig.rr_lam};
  for (DevBuf* b : bufs) b->release();
  h->batch_tot.release();
  h->call_sum.release();
  for (DevBuf* b : {&h->b_order, &h->b_blk, &h->b_seg, &h->b_q, &h->b_stats, &h->b_ticks,
                    &h->sp_xb, &h->sp_xb32, &h->sp_axb, &h->sp_nan})
    b->release();
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->g_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return DIDA_OK;
}

int dida_letkf_set_stream(dida_letkf_handle* h, void* s) {
  if (!h) return set_error(DIDA_ERR_ARG, "null handle");
  h->stream = s ? (hipStream_t)s : h->own_stream;
  return DIDA_OK;
}

int dida_letkf_set_option(dida_letkf_handle* h, int option, long long value) {
  if (!h) return set_error(DIDA_ERR_ARG, "null handle");
  switch (option) {
    case DIDA_OPT_SCHEDULE:
      if (value != 0 && value != 1) break;
      h->opt_schedule = (int)value;
      return DIDA_OK;
    case DIDA_OPT_TIMING:
      if (value != 0 && value != 1) break;
      h->opt_timing = (int)value;
      return DIDA_OK;
    case DIDA_OPT_TUNE:
      if (value != 0 && value != 1) break;
      h->opt_tune = (int)value;
      return DIDA_OK;
    case DIDA_OPT_GRAPH:
      if (value != 0 && value != 1) break;
      h->opt_graph = (int)value;
      return DIDA_OK;
    case DIDA_OPT_ROW_CAP:
      if (value < 0) break;
      h->row_cap = value;
      return DIDA_OK;
    default:
      return set_error(DIDA_ERR_ARG, "unknown option %d", option);
  }
  return set_error(DIDA_ERR_ARG, "bad value %lld for option %d", value, option);
}

int dida_letkf_set_grid(dida_letkf_handle* h, const float* lat, const float* lon) {
  if (!h || !lat || !lon) return set_error(DIDA_ERR_ARG, "null argument");
  const auto& c = h->cfg;
  RC_TRY(h->glat.ensure(sizeof(float) * c.nlat));
  RC_TRY(h->glon.ensure(sizeof(float) * c.nlon));
  HIP_TRY(hipMemcpyAsync(h->glat.p, lat, sizeof(float) * c.nlat, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(h->glon.p, lon, sizeof(float) * c.nlon, hipMemcpyHostToDevice, h->stream));
  if (!h->hyai.empty()) {
    RC_TRY(h->d_hyai.ensure(sizeof(double) * h->hyai.size()));
    RC_TRY(h->d_hybi.ensure(sizeof(double) * h->hybi.size()));
    HIP_TRY(hipMemcpyAsync(h->d_hyai.p, h->hyai.data(), sizeof(double) * h->hyai.size(),
                           hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->d_hybi.p, h->hybi.data(), sizeof(double) * h->hybi.size(),
                           hipMemcpyHostToDevice, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->have_grid = true;
  return DIDA_OK;
}

// Builds the obs index from device arrays (raw layout as the C ABI).
static int build_obs_index(dida_letkf_handle* h, int n_sites, const double* lat,
                           const double* lon, const int* row_ptr, int n_rows, const double* yb,
                           const double* dep, const double* rinv, const double* vco) {
  const auto& c = h->cfg;
  hipStream_t st = h->stream;
  h->n_sites = n_sites;
  h->n_rows = n_rows;
  const int nbins = h->nb_lat * h->nb_lon;
  RC_TRY(h->key_in.ensure(sizeof(int) * std::max(n_sites, 1)));
  RC_TRY(h->key_out.ensure(sizeof(int) * std::max(n_sites, 1)));
  RC_TRY(h->val_in.ensure(sizeof(int) * std::max(n_sites, 1)));
  RC_TRY(h->val_out.ensure(sizeof(int) * std::max(n_sites, 1)));
  RC_TRY(h->hist.ensure(sizeof(int) * (nbins + 1)));
  RC_TRY(h->bin_start.ensure(sizeof(int) * (nbins + 1)));
  HIP_TRY(hipMemsetAsync(h->hist.p, 0, sizeof(int) * (nbins + 1), st));
  HIP_TRY(launch_site_keys(st, n_sites, lat, lon, c.obs_polar, h->nb_lat, h->nb_lon, h->bdeg_lat,
                           h->bdeg_lon, h->key_in.as<int>(), h->val_in.as<int>(),
                           h->hist.as<int>()));
  // stable sort of site indices by bin key (ties keep input order)
  int end_bit = 1;
  while ((1 << end_bit) <= nbins) ++end_bit;
  size_t tmp_bytes = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, h->key_in.as<int>(),
                                             h->key_out.as<int>(), h->val_in.as<int>(),
                                             h->val_out.as<int>(), n_sites, 0, end_bit, st));
  size_t scan_bytes = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, h->hist.as<int>(),
                                           h->bin_start.as<int>(), nbins + 1, st));
  RC_TRY(h->sort_tmp.ensure(std::max(tmp_bytes, scan_bytes)));
  if (n_sites > 0)
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(h->sort_tmp.p, tmp_bytes, h->key_in.as<int>(),
                                               h->key_out.as<int>(), h->val_in.as<int>(),
                                               h->val_out.as<int>(), n_sites, 0, end_bit, st));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(h->sort_tmp.p, scan_bytes, h->hist.as<int>(),
                                           h->bin_start.as<int>(), nbins + 1, st));
  const size_t ns = std::max(n_sites, 1);
  RC_TRY(h->s_cos.ensure(sizeof(double) * ns));
  RC_TRY(h->s_sin.ensure(sizeof(double) * ns));
  RC_TRY(h->s_lonr.ensure(sizeof(double) * ns));
  RC_TRY(h->s_cosl.ensure(sizeof(double) * ns));
  RC_TRY(h->s_sinl.ensure(sizeof(double) * ns));
  RC_TRY(h->s_row0.ensure(sizeof(int) * ns));
  RC_TRY(h->s_nrow.ensure(sizeof(int) * ns));
  HIP_TRY(launch_site_gather(st, n_sites, h->val_out.as<int>(), lat, lon, row_ptr,
                             h->s_cos.as<double>(), h->s_sin.as<double>(), h->s_lonr.as<double>(),
                             h->s_cosl.as<double>(), h->s_sinl.as<double>(), h->s_row0.as<int>(),
                             h->s_nrow.as<int>()));
  // rows in the storage precision of the mode
  const long long nr = std::max(n_rows, 1);
  RC_TRY(h->rinv32.ensure(sizeof(float) * nr));
  if (n_rows > 0)
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, st,
                       (long long)n_rows, rinv, h->rinv32.as<float>());
  if (c.precision != DIDA_PREC_FP64) {
    const long long ny = (long long)n_rows * c.k;
    // the reduced-precision products address the FP32 Yb with 32-bit byte
    // offsets below kYbOffsetLimit
    if (4 * (ny + (long long)kListPad * c.k) >= kYbOffsetLimit)
      return set_error(DIDA_ERR_UNSUPPORTED, "%d obs rows x k=%d exceed the 2 GiB of FP32 Yb "
                       "the reduced-precision products address", n_rows, c.k);
    RC_TRY(h->yb32.ensure(sizeof(float) * (std::max(ny, 1LL) + (long long)kListPad * c.k)));
    if (ny > 0)
      hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, st, ny,
                         yb, h->yb32.as<float>());
  }
  HIP_TRY(hipGetLastError());
  h->have_obs = true;
  return DIDA_OK;
}

static int stage_obs(dida_letkf_handle* h, int n_sites, const double* lat, const double* lon,
                     const int* row_ptr, int n_rows, const double* yb, const double* dep,
                     const double* rinv, const double* vco, hipMemcpyKind kind) {
  const auto& c = h->cfg;
  hipStream_t st = h->stream;
  const size_t ns = std::max(n_sites, 1), nr = std::max(n_rows, 1);
  RC_TRY(h->o_lat.ensure(sizeof(double) * ns));
  RC_TRY(h->o_lon.ensure(sizeof(double) * ns));
  RC_TRY(h->o_yb64.ensure(sizeof(double) * nr * c.k));
  RC_TRY(h->o_dep.ensure(sizeof(double) * nr));
  RC_TRY(h->o_rinv64.ensure(sizeof(double) * nr));
  HIP_TRY(hipMemcpyAsync(h->o_lat.p, lat, sizeof(double) * n_sites, kind, st));
  HIP_TRY(hipMemcpyAsync(h->o_lon.p, lon, sizeof(double) * n_sites, kind, st));
  HIP_TRY(hipMemcpyAsync(h->o_yb64.p, yb, sizeof(double) * (size_t)n_rows * c.k, kind, st));
  HIP_TRY(hipMemcpyAsync(h->o_dep.p, dep, sizeof(double) * n_rows, kind, st));
  HIP_TRY(hipMemcpyAsync(h->o_rinv64.p, rinv, sizeof(double) * n_rows, kind, st));
  const int* d_rowptr = nullptr;
  if (row_ptr) {
    RC_TRY(h->o_rowptr.ensure(sizeof(int) * (ns + 1)));
    HIP_TRY(hipMemcpyAsync(h->o_rowptr.p, row_ptr, sizeof(int) * (n_sites + 1), kind, st));
    d_rowptr = h->o_rowptr.as<int>();
  }
  if (vco) {
    RC_TRY(h->o_vco.ensure(sizeof(double) * nr));
    HIP_TRY(hipMemcpyAsync(h->o_vco.p, vco, sizeof(double) * n_rows, kind, st));
  } else if (c.vert_dist > 0.0) {
    return set_error(DIDA_ERR_ARG, "vertical localisation needs vcoord");
  }
  return build_obs_index(h, n_sites, h->o_lat.as<double>(), h->o_lon.as<double>(), d_rowptr,
                         n_rows, h->o_yb64.as<double>(), h->o_dep.as<double>(),
                         h->o_rinv64.as<double>(), vco ? h->o_vco.as<double>() : nullptr);
}

static int check_obs_args(dida_letkf_handle* h, int n_sites, const double* lat, const double* lon,
                          const int* row_ptr, int n_rows, const double* yb, const double* dep,
                          const double* rinv, bool host) {
  if (!h) return set_error(DIDA_ERR_ARG, "null handle");
  if (n_sites < 0 || n_rows < 0) return set_error(DIDA_ERR_ARG, "negative sizes");
  if (n_sites > 0 && (!lat || !lon)) return set_error(DIDA_ERR_ARG, "null site arrays");
  if (n_rows > 0 && (!yb || !dep || !rinv)) return set_error(DIDA_ERR_ARG, "null row arrays");
  if (!row_ptr && n_rows != n_sites)
    return set_error(DIDA_ERR_ARG, "one row per site needs n_rows == n_sites");
  if (row_ptr && host) {
    if (row_ptr[0] != 0 || row_ptr[n_sites] != n_rows)
      return set_error(DIDA_ERR_ARG, "site_row_ptr must run from 0 to n_rows");
    for (int s = 0; s < n_sites; ++s)
      if (row_ptr[s + 1] < row_ptr[s]) return set_error(DIDA_ERR_ARG, "site_row_ptr decreasing");
  }
  return DIDA_OK;
}

int dida_letkf_set_obs(dida_letkf_handle* h, int n_sites, const double* lat, const double* lon,
                       const int* row_ptr, int n_rows, const double* yb, const double* dep,
                       const double* rinv, const double* vco) {
  RC_TRY(check_obs_args(h, n_sites, lat, lon, row_ptr, n_rows, yb, dep, rinv, true));
  RC_TRY(stage_obs(h, n_sites, lat, lon, row_ptr, n_rows, yb, dep, rinv, vco,
                   hipMemcpyHostToDevice));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return DIDA_OK;
}

int dida_letkf_set_obs_device(dida_letkf_handle* h, int n_sites, const double* lat,
                              const double* lon, const int* row_ptr, int n_rows,
                              const double* yb, const double* dep, const double* rinv,
                              const double* vco) {
  RC_TRY(check_obs_args(h, n_sites, lat, lon, row_ptr, n_rows, yb, dep, rinv, false));
  return stage_obs(h, n_sites, lat, lon, row_ptr, n_rows, yb, dep, rinv, vco,
                   hipMemcpyDeviceToDevice);
}

}  // extern "C"

// Shared analysis driver.  xb etc. are device pointers.
// The state an analysis call reads and writes: FP64 or FP32 perturbations,
// FP64 means, and either xa/axa or the fused caller epilogue's outputs.
struct StateArgs {
  const double* xb = nullptr;
  const float* xb32 = nullptr;
  const double* axb = nullptr;
  double* xa = nullptr;
  float* xa32 = nullptr;
  double* axa = nullptr;
  int epi = 0;
  double polar_lat = 90.0, gamma1 = 0.0, gamma2 = 1.0;
  double* xa_out = nullptr;
  double* spread_b = nullptr;
  double* spread_a = nullptr;
  // the caller's full ensemble state, split on the device into xb / xb32 and
  // axb (handle buffers) before the analysis
  const double* x_in = nullptr;
};

// ---------------------------------------------------------------------------
// Tuning cache: the products variant (32-row groups per step) of each
// (device, k, precision, length bucket), measured on a live call and kept in
// a text file, one "arch k precision bucket variant microseconds" line each.

namespace {
std::mutex g_tune_mu;
std::map<std::string, std::pair<int, double>> g_tune;
bool g_tune_loaded = false;

std::string tune_path() {
  if (const char* f = std::getenv("DIDA_TUNE_FILE")) return f;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) return std::string(x) + "/dida_letkf/tune.tsv";
  if (const char* hm = std::getenv("HOME")) return std::string(hm) + "/.cache/dida_letkf/tune.tsv";
  return std::string();
}

std::string tune_key(const std::string& arch, int k, int prec, int bucket) {
  return arch + " " + std::to_string(k) + " " + std::to_string(prec) + " " + std::to_string(bucket);
}

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const std::string path = tune_path();
  if (path.empty()) return;
  std::ifstream f(path);
  std::string arch;
  int k, prec, bucket, variant;
  double us;
  while (f >> arch >> k >> prec >> bucket >> variant >> us)
    if (variant >= 1 && variant <= 4) g_tune[tune_key(arch, k, prec, bucket)] = {variant, us};
}

// rewrites the whole file (entries of every device and shape seen)
int tune_save_locked() {
  const std::string path = tune_path();
  if (path.empty()) return DIDA_OK;
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  {
    std::ofstream f(tmp);
    if (!f) return set_error(DIDA_ERR_ARG, "cannot write the tuning cache %s", tmp.c_str());
    for (const auto& kv : g_tune) f << kv.first << " " << kv.second.first << " " << kv.second.second << "\n";
  }
  std::filesystem::rename(tmp, path, ec);
  if (ec) return set_error(DIDA_ERR_ARG, "cannot replace the tuning cache %s", path.c_str());
  return DIDA_OK;
}

// Products variants of a bucket: 1, 2 = dispatch order with 1 or 2 row
// groups of 32 per step, 3, 4 = the same on a persistent work queue.  Two row
// groups exist for BF16 at k <= 64.  Every variant gives the same bits.
std::vector<int> products_variants(int kp, int precision) {
  if (kp <= 64 && precision == DIDA_PREC_BF16X3) return {1, 2, 3, 4};
  return {1, 3};
}
int variant_groups(int v) { return v == 2 || v == 4 ? 2 : 1; }
bool variant_persistent(int v) { return v >= 3; }
}  // namespace

// variant per bucket position from the cache (default 1: dispatch order,
// one row group); DIDA_PRODUCTS_VARIANT = 1..4 sets every bucket's (a test hook)
static void resolve_variants(dida_letkf_handle* h) {
  const std::vector<int> vars = products_variants(h->kp, h->cfg.precision);
  const int forced = env_int("DIDA_PRODUCTS_VARIANT");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  for (int j = 0; j < kBuckets; ++j) {
    const int b = kBuckets - 1 - j;
    auto it = g_tune.find(tune_key(h->arch, h->cfg.k, h->cfg.precision, b));
    int v = it != g_tune.end() ? it->second.first : 1;
    if (forced) v = forced;
    h->variant[j] = std::find(vars.begin(), vars.end(), v) != vars.end() ? v : 1;
  }
}

// Row-list capacity per batch before any call has measured one: the expected
// rows of uniformly spread obs within the support radius 2c, twice over, plus
// a margin (a call that finds more runs again with the measured need).
static long long estimate_row_cap(const dida_letkf_handle* h) {
  const double ang = std::min(M_PI, 2.0 * h->cfg.local_dist / kEarthF32);
  const double frac = 0.5 * (1.0 - std::cos(ang));
  const double per_col = (double)h->n_rows * frac;
  const double lists = h->tiled ? (double)h->max_batch / kTile * 2.0 : (double)h->max_batch;
  const double est = 2.0 * per_col * lists + 65536.0;
  const double most = (double)h->n_rows * (double)h->max_batch + 1.0;
  return (long long)std::min(est, most);
}

static long long n_batches_of(const dida_letkf_handle* h) {
  return (h->n_h + h->max_batch - 1) / h->max_batch;
}

// Every buffer of a call, so that its launches allocate nothing.
static int prepare_call(dida_letkf_handle* h) {
  const auto& c = h->cfg;
  const long long mb = h->max_batch, nb = n_batches_of(h);
  if (h->row_cap <= 0) h->row_cap = estimate_row_cap(h);
  const long long rc = std::max(h->row_cap, 1LL);
  RC_TRY(h->cnt.ensure(sizeof(long long) * (mb + 1)));
  RC_TRY(h->off.ensure(sizeof(long long) * (mb + 1)));
  RC_TRY(h->scan_tmp.ensure(sizeof(long long) * scan_tmp_elems(mb + 1)));
  RC_TRY(h->Abuf.ensure(sizeof(double) * mb * h->nzc * h->kp * h->kp));
  RC_TRY(h->bbuf.ensure(sizeof(double) * mb * h->nzc * h->kp));
  RC_TRY(h->status.ensure(sizeof(int) * h->n_an));
  RC_TRY(h->nobsl_dev.ensure(sizeof(int) * std::max(h->n_h, 1LL)));
  RC_TRY(h->batch_tot.ensure(sizeof(long long) * nb));
  {
    // the lists' padding holds valid row indices: a new buffer is zeroed
    // (once, before any call's launches)
    const void* before = h->row_idx.p;
    RC_TRY(h->row_idx.ensure(sizeof(int) * (rc + kListPad)));
    if (h->row_idx.p != before)
      HIP_TRY(launch_zero_ints(h->stream, h->row_idx.as<int>(), (int)(rc + kListPad)));
  }
  if (h->tiled) {
    RC_TRY(h->row_site.ensure(sizeof(int) * rc));
  } else {
    RC_TRY(h->row_gc.ensure(sizeof(double) * (rc + kListPad)));
    if (h->nzc == 1) RC_TRY(h->row_wd.ensure(sizeof(double) * (rc + kListPad)));
  }
  if (c.keep_transforms || c.overlap > 0)
    RC_TRY(h->trans.ensure(sizeof(double) * h->n_an * c.k * c.k));
  if (c.keep_transforms) {
    RC_TRY(h->wout.ensure(sizeof(double) * h->n_an * c.k));
    h->nobsl_host.assign(h->n_an, 0);
  }
  RC_TRY(eig_prepare(h->eig, c.k, mb * h->nzc));
  if (c.transform_kernel == 0 && mb * h->nzc >= 2 * kEigPipeMinHalf)
    RC_TRY(eig_second_stream(h->eig));
  if (h->opt_schedule == 1 && !h->tiled) {
    RC_TRY(h->b_order.ensure(sizeof(int) * mb * h->nzc));
    RC_TRY(h->b_blk.ensure(sizeof(long long) * std::max(bucket_blk_elems((int)mb), 1LL)));
    RC_TRY(h->b_seg.ensure(sizeof(long long) * (kBuckets + 1)));
    RC_TRY(h->b_q.ensure(sizeof(int) * kBuckets * kQueueParts));
    RC_TRY(h->b_stats.ensure(sizeof(long long) * nb * 2 * kBuckets));
    RC_TRY(h->b_ticks.ensure(sizeof(unsigned long long) * nb * 5 * kBuckets));
  }
  while ((long long)h->ev.size() < 9 * nb) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    h->ev.push_back(e);
  }
  return DIDA_OK;
}

// Measures each products variant on every non-empty bucket of the batch in
// pp (buckets already built), keeps the fastest per bucket in the handle and
// the cache, and leaves the queue counters zeroed.  Eager calls only.
static int tune_products(dida_letkf_handle* h, hipStream_t st, ProdParams pp) {
  const auto& c = h->cfg;
  long long seg[kBuckets + 1];
  HIP_TRY(hipMemcpyAsync(seg, h->b_seg.p, sizeof(seg), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const std::vector<int> vars = products_variants(h->kp, c.precision);
  for (auto& e : h->g_ev)
    if (!e) HIP_TRY(hipEventCreate(&e));
  hipEvent_t e0 = h->g_ev[0], e1 = h->g_ev[1];
  double best_us[kBuckets];
  int best[kBuckets];
  for (int j = 0; j < kBuckets; ++j) {
    best_us[j] = 1e300;
    best[j] = h->variant[j];
  }
  pp.bucket_ticks = nullptr;
  pp.launch_ticks = nullptr;
  for (int rep = 0; rep < 2; ++rep) {
    for (int g : vars) {
      for (int j = 0; j < kBuckets; ++j) {
        if (seg[j + 1] <= seg[j]) continue;
        pp.seg_lo = j;
        pp.seg_hi = j + 1;
        pp.q_counter = variant_persistent(g) ? h->b_q.as<int>() + j * kQueueParts : nullptr;
        pp.groups = variant_groups(g);
        if (pp.q_counter) HIP_TRY(launch_zero_ints(st, pp.q_counter, kQueueParts));
        HIP_TRY(hipEventRecord(e0, st));
        HIP_TRY(launch_products_queue(st, pp, c.precision));
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (1000.0 * ms < best_us[j]) {
          best_us[j] = 1000.0 * ms;
          best[j] = g;
        }
      }
    }
  }
  HIP_TRY(launch_zero_ints(st, h->b_q.as<int>(), kBuckets * kQueueParts));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  for (int j = 0; j < kBuckets; ++j) {
    if (seg[j + 1] <= seg[j]) continue;
    h->variant[j] = best[j];
    g_tune[tune_key(h->arch, c.k, c.precision, kBuckets - 1 - j)] = {best[j], best_us[j]};
  }
  return tune_save_locked();
}

// The launches of one analysis call, every batch in turn on the handle's
// stream: no allocation and no host synchronisation (unless a tuning pass
// runs), so that a call can be captured in a hipGraph.  timed: record the
// per-batch events; tmode: run_transform's mode.
static int enqueue_call(dida_letkf_handle* h, const StateArgs& sa, bool timed, int tmode,
                        bool tune) {
  const double* axb = sa.axb;
  const auto& c = h->cfg;
  hipStream_t st = h->stream;
  Geom g;
  g.glat = h->glat.as<float>();
  g.glon = h->glon.as<float>();
  g.nlat = c.nlat;
  g.nlon = c.nlon;
  g.cube_n = c.cube_n;
  g.cube_s = c.overlap > 0 ? c.cube_n - c.overlap : c.cube_n;
  g.nlat_c = h->nlat_c;
  g.nlon_c = h->nlon_c;
  const long long mb = h->max_batch;
  const bool tiled = h->tiled;
  const bool queue = h->opt_schedule == 1 && !tiled;
  HIP_TRY(launch_zero_ints(st, h->eig.flags.as<int>(), 2));
  if (sa.x_in) {
    SplitParams spp{};
    spp.x_in = sa.x_in;
    spp.xb = const_cast<double*>(sa.xb);
    spp.xb32 = const_cast<float*>(sa.xb32);
    spp.axb = const_cast<double*>(sa.axb);
    spp.n_elems = (long long)(c.nlat + 2 * c.halo) * (c.nlon + 2 * c.halo) * h->n;
    spp.k = c.k;
    spp.nan = h->sp_nan.as<unsigned long long>();
    HIP_TRY(launch_split(st, spp));
  }

  long long bi = 0;
  for (long long h0 = 0; h0 < h->n_h; h0 += mb, ++bi) {
    const long long h1 = std::min(h->n_h, h0 + mb);
    const int ncol = (int)(h1 - h0);
    hipEvent_t* ev = h->ev.data() + 9 * bi;
    if (timed) HIP_TRY(hipEventRecord(ev[0], st));
    SelectParams sp{};
    sp.g = g;
    sp.s_cos = h->s_cos.as<double>();
    sp.s_sin = h->s_sin.as<double>();
    sp.s_lonr = h->s_lonr.as<double>();
    sp.s_cosl = h->s_cosl.as<double>();
    sp.s_sinl = h->s_sinl.as<double>();
    {
      // the support test ri <= 2 in cos(angle): v >= cos(2c / a) (every
      // point when 2c / a >= pi), with a guard band far wider than the
      // rounding of either evaluation
      const double ang = 2.0 * c.local_dist / kEarthF32;
      const double thr = ang >= M_PI ? -1.0 : std::cos(ang);
      sp.v_lo = thr - 1e-12;
      sp.v_hi = thr + 1e-12;
    }
    sp.s_row0 = h->s_row0.as<int>();
    sp.s_nrow = h->s_nrow.as<int>();
    sp.bin_start = h->bin_start.as<int>();
    sp.nb_lat = h->nb_lat;
    sp.nb_lon = h->nb_lon;
    sp.bdeg_lat = h->bdeg_lat;
    sp.bdeg_lon = h->bdeg_lon;
    sp.theta_deg = h->theta_deg;
    sp.local_dist = c.local_dist;
    sp.gc_filter = c.gc_filter;
    sp.h0 = (int)h0;
    sp.h1 = (int)h1;
    sp.cnt = h->cnt.as<long long>();
    sp.off = h->off.as<long long>();
    sp.tile = tiled ? kTile : 0;
    if (products_band() > 0 && h0 % g.nlat_c == 0 && (h1 - h0) % g.nlat_c == 0) {
      sp.band_w = products_band();
      sp.band_l = g.nlat_c;
    }
    // horizontal loop: weights precomputed by the fill pass (for the BF16
    // products as the FP32 pair {sqrt(w), w dep})
    sp.weights = h->nzc != 1 ? 0 : c.precision == DIDA_PREC_BF16X3 ? 2 : 1;
    sp.k = c.k;
    sp.rinv32 = h->rinv32.as<float>();
    sp.dep = h->o_dep.as<double>();
    // count pass -> offsets -> fill pass into the fixed-capacity lists; per
    // analysis (column mode) or per tile of kTile analyses (tile mode: the
    // union of their candidate rows).  The list total goes to batch_tot and
    // the rows per column to nobsl_dev (tile mode: from the products kernel).
    const int nlist = tiled ? (ncol + kTile - 1) / kTile : ncol;
    HIP_TRY(launch_zero_ll(st, h->cnt.as<long long>(), nlist + 1));
    if (tiled) HIP_TRY(launch_select_tile(st, sp, false));
    else HIP_TRY(launch_select(st, sp, false));
    HIP_TRY(launch_exclusive_scan(st, h->cnt.as<long long>(), h->off.as<long long>(), nlist + 1,
                                  h->scan_tmp.as<long long>()));
    HIP_TRY(launch_list_counts(st, h->off.as<long long>(), nlist,
                               tiled ? nullptr : h->nobsl_dev.as<int>() + h0,
                               h->batch_tot.as<long long>() + bi));
    sp.row_idx = h->row_idx.as<int>();
    sp.cap = h->row_cap;
    if (tiled) {
      sp.row_site = h->row_site.as<int>();
      HIP_TRY(launch_select_tile(st, sp, true));
    } else {
      sp.row_gc = h->row_gc.as<double>();
      sp.row_wd = sp.weights ? h->row_wd.as<double>() : nullptr;
      HIP_TRY(launch_select(st, sp, true));
    }
    if (timed) HIP_TRY(hipEventRecord(ev[1], st));

    ProdParams pp{};
    pp.g = g;
    pp.k = c.k;
    pp.kp = h->kp;
    pp.nzc = h->nzc;
    pp.vertical = c.vert_dist > 0.0;
    pp.vcoord_mode = c.vcoord_mode;
    pp.cube_nz = h->cfg.cube_nz;
    pp.nz = c.nz;
    pp.halo = c.halo;
    pp.n = h->n;
    pp.ps_index = c.ps_index;
    pp.vert_dist = c.vert_dist;
    pp.gc_filter_vertical = c.gc_filter_vertical;
    pp.hyai = h->d_hyai.as<double>();
    pp.hybi = h->d_hybi.as<double>();
    pp.p0 = c.p0;
    pp.ptop = c.ptop;
    pp.axb = axb;
    pp.row_idx = h->row_idx.as<int>();
    pp.row_gc = h->row_gc.as<double>();
    pp.row_wd = h->row_wd.as<double>();
    pp.pre_weighted = (!tiled && h->nzc == 1) ? 1 : 0;
    pp.row_off = h->off.as<long long>();
    pp.row_cap = h->row_cap;
    pp.yb64 = h->o_yb64.as<double>();
    pp.yb32 = h->yb32.as<float>();
    pp.yb32_bytes = (unsigned)std::min<size_t>(h->yb32.bytes, kYbOffsetLimit);
    pp.dep = h->o_dep.as<double>();
    pp.rinv32 = h->rinv32.as<float>();
    pp.vco = h->o_vco.as<double>();
    pp.a0 = h0 * h->nzc;
    pp.n_an = (h1 - h0) * h->nzc;
    pp.groups = 1;
    pp.h0 = (int)h0;
    pp.A = h->Abuf.as<double>();
    pp.b = h->bbuf.as<double>();
    if (tiled) {
      pp.tile = kTile;
      pp.row_site = h->row_site.as<int>();
      pp.s_cos = h->s_cos.as<double>();
      pp.s_sin = h->s_sin.as<double>();
      pp.s_lonr = h->s_lonr.as<double>();
      pp.local_dist = c.local_dist;
      pp.gc_filter = c.gc_filter;
      pp.nobsl = h->nobsl_dev.as<int>() + h0;
      HIP_TRY(launch_products_tile(st, pp, c.precision));
    } else if (queue) {
      BucketParams bp{};
      bp.counts = h->nobsl_dev.as<int>() + h0;
      bp.n_cols = ncol;
      bp.nzc = h->nzc;
      bp.blk = h->b_blk.as<long long>();
      bp.seg_off = h->b_seg.as<long long>();
      bp.bucket_stats = h->b_stats.as<long long>() + bi * 2 * kBuckets;
      bp.order = h->b_order.as<int>();
      bp.q_counters = h->b_q.as<int>();
      bp.ticks = h->opt_timing ? h->b_ticks.as<unsigned long long>() + bi * 5 * kBuckets : nullptr;
      if (products_band() > 0 && h0 % g.nlat_c == 0 && (h1 - h0) % g.nlat_c == 0) {
        bp.band_w = products_band();
        bp.band_l = g.nlat_c;
      }
      HIP_TRY(launch_buckets(st, bp));
      pp.order = bp.order;
      pp.seg_off = bp.seg_off;
      if (tune && bi == 0 && products_variants(h->kp, c.precision).size() > 1)
        RC_TRY(tune_products(h, st, pp));
      // runs of bucket positions that share a variant: one launch each
      for (int j0 = 0; j0 < kBuckets;) {
        int j1 = j0 + 1;
        while (j1 < kBuckets && h->variant[j1] == h->variant[j0]) ++j1;
        const int v = h->variant[j0];
        pp.seg_lo = j0;
        pp.seg_hi = j1;
        pp.q_counter = variant_persistent(v) ? bp.q_counters + j0 * kQueueParts : nullptr;
        pp.groups = variant_groups(v);
        pp.bucket_ticks = bp.ticks;
        pp.launch_ticks = bp.ticks ? bp.ticks + kBuckets + 4 * j0 : nullptr;
        long long cap = 0;
        HIP_TRY(launch_products_queue(st, pp, c.precision, &cap));
        if (timed && (size_t)(bi * kBuckets + j0) < h->run_capacity.size())
          h->run_capacity[bi * kBuckets + j0] = cap;
        j0 = j1;
      }
    } else {
      if (products_band() > 0 && h->nzc == 1 && h0 % g.nlat_c == 0 &&
          (h1 - h0) % g.nlat_c == 0) {
        pp.band_w = products_band();
        pp.band_l = g.nlat_c;
      }
      HIP_TRY(launch_products(st, pp, c.precision));
    }
    if (timed) HIP_TRY(hipEventRecord(ev[2], st));

    XformParams xp{};
    xp.g = g;
    xp.k = c.k;
    xp.kp = h->kp;
    // the BF16 products of k <= 64 store A in FP32 tile order
    xp.a_f32 = !tiled && c.precision == DIDA_PREC_BF16X3 && h->kp <= 64;
    xp.inflation = c.inflation_factor;
    xp.guard = c.precision != DIDA_PREC_FP64;
    xp.A = pp.A;
    xp.b = pp.b;
    xp.a0 = pp.a0;
    xp.n_an = pp.n_an;
    xp.nzc = h->nzc;
    xp.vertical = pp.vertical;
    xp.cube_nz = h->cfg.cube_nz;
    xp.nz = c.nz;
    xp.halo = c.halo;
    xp.n = h->n;
    xp.n_var = c.n2d + c.n3d;
    {
      int cur = 1;
      for (int v = 0; v < c.n2d; ++v) { xp.var_start[v] = cur; xp.var_lev[v] = 1; cur += 1; }
      for (int v = 0; v < c.n3d; ++v) {
        xp.var_start[c.n2d + v] = cur;
        xp.var_lev[c.n2d + v] = c.nz;
        cur += c.nz;
      }
    }
    // overlapped cubes: the eigensolver writes each cube's trans only, and
    // k_overlap_update forms every column from its cubes after the last batch
    const bool ovl = c.overlap > 0;
    xp.xb = ovl ? nullptr : sa.xb;
    xp.axb = axb;
    xp.xa = ovl ? nullptr : sa.xa;
    xp.axa = ovl ? nullptr : sa.axa;
    xp.state_f32 = sa.xb32 != nullptr;
    xp.xb32 = sa.xb32;
    xp.xa32 = sa.xa32;
    xp.epi = sa.epi;
    xp.polar_lat = sa.polar_lat;
    xp.gamma1 = sa.gamma1;
    xp.gamma2 = sa.gamma2;
    xp.xa_out = sa.xa_out;
    xp.spread_b = sa.spread_b;
    xp.spread_a = sa.spread_a;
    xp.trans_out = (c.keep_transforms || ovl) ? h->trans.as<double>() : nullptr;
    xp.w_out = c.keep_transforms ? h->wout.as<double>() : nullptr;
    xp.status = h->status.as<int>();
    xp.list = nullptr;
    bool split = false;
    RC_TRY(run_transform(st, h->eig, xp, tmode, timed ? ev + 4 : nullptr, &split));
    h->split_timed[bi] = timed && split;
    if (ovl && h1 == h->n_h) {
      OverlapParams op{};
      op.g = g;
      op.k = c.k;
      op.n = h->n;
      op.halo = c.halo;
      op.xb = sa.xb;
      op.axb = axb;
      op.xa = sa.xa;
      op.axa = sa.axa;
      op.trans = h->trans.as<double>();
      op.status = h->status.as<int>();
      HIP_TRY(launch_overlap_update(st, op));
    }
    if (timed) HIP_TRY(hipEventRecord(ev[3], st));
  }
  return DIDA_OK;
}

// The reference stops on a NaN background member (module_da_parallel.F90:
// 189-193); here the call fails with the first such element, in the
// reference's (k, j, i) indices.
static int nan_error(const dida_letkf_handle* h, const unsigned long long* nan) {
  const auto& c = h->cfg;
  const long long n_elems = (long long)(c.nlat + 2 * c.halo) * (c.nlon + 2 * c.halo) * h->n;
  const long long el = n_elems - (long long)nan[1];
  const long long col = el / h->n, lat_ext = c.nlat + 2 * c.halo;
  return set_error(DIDA_ERR_ARG,
                   "background NaN in %llu state elements; first at k=%lld j=%lld i=%lld",
                   nan[0], el % h->n + 1, col % lat_ext + 1 - c.halo,
                   col / lat_ext + 1 - c.halo);
}

// After the call's launches: one readback (row totals and counts, status,
// flags, bucket statistics), the stats, and the largest row total of a batch
// in need (the call's lists were cut if it exceeds row_cap).  Returns the
// number of analyses still flagged.
static int finish_call(dida_letkf_handle* h, bool timed, float graph_ms, bool split,
                       dida_letkf_stats* stats, long long& need) {
  const auto& c = h->cfg;
  hipStream_t st = h->stream;
  const long long nb = n_batches_of(h);
  const bool queue = h->opt_schedule == 1 && !h->tiled;
  const int n_bst = queue ? (int)(nb * 2 * kBuckets) : 0;
  const int n_tk = queue && h->opt_timing ? (int)(nb * 5 * kBuckets) : 0;
  const size_t words = kSumHead + (size_t)nb + n_bst + n_tk;
  RC_TRY(h->call_sum.ensure(sizeof(long long) * words));
  SummaryParams sp{};
  sp.status = h->status.as<int>();
  sp.n_an = h->n_an;
  sp.counts = h->nobsl_dev.as<int>();
  sp.n_h = h->n_h;
  sp.nzc = h->nzc;
  sp.flags = h->eig.flags.as<int>();
  sp.nan = split ? h->sp_nan.as<unsigned long long>() : nullptr;
  sp.tot = h->batch_tot.as<long long>();
  sp.n_tot = (int)nb;
  sp.bst = n_bst ? h->b_stats.as<long long>() : nullptr;
  sp.n_bst = n_bst;
  sp.tk = n_tk ? h->b_ticks.as<unsigned long long>() : nullptr;
  sp.n_tk = n_tk;
  sp.out = h->call_sum.as<long long>();
  hipLaunchKernelGGL(k_call_summary, dim3(1), dim3(1024), 0, st, sp);
  HIP_TRY(hipGetLastError());
  std::vector<long long> sum(words);
  HIP_TRY(hipMemcpyAsync(sum.data(), h->call_sum.p, sizeof(long long) * words,
                         hipMemcpyDeviceToHost, st));
  std::vector<int> counts;
  if (c.keep_transforms) {
    counts.resize(h->n_h);
    HIP_TRY(hipMemcpyAsync(counts.data(), h->nobsl_dev.p, sizeof(int) * h->n_h,
                           hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  unsigned long long nan[2] = {(unsigned long long)sum[6], (unsigned long long)sum[7]};
  if (nan[0]) return nan_error(h, nan);
  const long long* tot = sum.data() + kSumHead;
  const long long* bst = tot + nb;
  const long long* tk = bst + n_bst;
  need = 0;
  for (long long b = 0; b < nb; ++b) need = std::max(need, tot[b]);
  if (need > h->row_cap) return 0;
  const int flagged = (int)sum[0];
  if (c.keep_transforms)
    for (long long a = 0; a < h->n_an; ++a) h->nobsl_host[a] = counts[a / h->nzc];
  if (!stats) return flagged;
  const long long total_rows = sum[1];
  const int max_rows = (int)sum[2], over_cap = (int)sum[3];
  float ms_sel = 0, ms_prod = 0, ms_xf = 0, ms_k[4] = {0, 0, 0, 0};
  if (timed) {
    for (long long b = 0; b < nb; ++b) {
      hipEvent_t* ev = h->ev.data() + 9 * b;
      float t1 = 0, t2 = 0, t3 = 0;
      HIP_TRY(hipEventElapsedTime(&t1, ev[0], ev[1]));
      HIP_TRY(hipEventElapsedTime(&t2, ev[1], ev[2]));
      HIP_TRY(hipEventElapsedTime(&t3, ev[2], ev[3]));
      ms_sel += t1;
      ms_prod += t2;
      ms_xf += t3;
      if (h->split_timed[b]) {
        for (int i = 0; i < 4; ++i) {
          float t = 0;
          HIP_TRY(hipEventElapsedTime(&t, ev[4 + i], ev[5 + i]));
          ms_k[i] += t;
        }
      }
    }
  }
  *stats = dida_letkf_stats{};
  stats->n_analyses = h->n_an;
  stats->total_rows = total_rows;
  stats->max_rows = max_rows;
  stats->n_flagged = flagged;
  stats->rows_over_10000 = over_cap;
  stats->ms_select = ms_sel;
  stats->ms_products = ms_prod;
  stats->ms_transform = ms_xf;
  stats->ms_total = timed ? ms_sel + ms_prod + ms_xf : graph_ms;
  stats->ms_tridiag = ms_k[0];
  stats->ms_ql = ms_k[1];
  stats->ms_apply = ms_k[2];
  stats->ms_update = ms_k[3];
  stats->n_rerun = (int)sum[5];
  stats->schedule = queue ? 1 : 0;
  if (queue) {
    double busy = 0, cap = 0;
    for (long long b = 0; b < nb; ++b) {
      for (int i = 0; i < kBuckets; ++i) {
        stats->bucket_analyses[i] += bst[b * 2 * kBuckets + i];
        stats->bucket_rows[i] += bst[b * 2 * kBuckets + kBuckets + i];
      }
      if (n_tk == 0) continue;
      const long long* t = tk + b * 5 * kBuckets;
      for (int j = 0; j < kBuckets; ++j)
        stats->bucket_wave_ms[kBuckets - 1 - j] += (double)t[j] / h->clock_khz;
      for (int j = 0; j < kBuckets; ++j) {
        const long long* l = t + kBuckets + 4 * j;
        const long long slots = h->run_capacity[b * kBuckets + j];
        if (l[3] == 0 || l[2] == 0 || slots <= 0) continue;
        busy += (double)l[2];
        cap += (double)slots * (double)(l[1] - l[0]);
      }
    }
    stats->products_balance = cap > 0 ? busy / cap : 0.0;
  }
  for (int j = 0; j < kBuckets; ++j) stats->variant[j] = h->variant[kBuckets - 1 - j];
  return flagged;
}

// Graph key: everything a captured call holds by value (pointers, sizes,
// options), compared before each replay.
static std::vector<uintptr_t> graph_key(const dida_letkf_handle* h, const StateArgs& sa) {
  std::vector<uintptr_t> k;
  auto put = [&](const void* p) { k.push_back((uintptr_t)p); };
  auto putd = [&](double d) {
    uintptr_t u = 0;
    std::memcpy(&u, &d, sizeof(d));
    k.push_back(u);
  };
  put(h->stream);
  // grid and obs buffers (set_obs refills them in place while they are large
  // enough, so a new cycle's obs replay the same graph)
  for (const DevBuf* b : {&h->glat, &h->glon, &h->o_yb64, &h->o_dep, &h->o_vco, &h->bin_start,
                          &h->s_cos, &h->s_sin, &h->s_lonr, &h->s_cosl, &h->s_sinl, &h->s_row0,
                          &h->s_nrow, &h->yb32, &h->rinv32, &h->d_hyai, &h->d_hybi})
    put(b->p);
  k.push_back((uintptr_t)h->row_cap);
  k.push_back((uintptr_t)h->opt_schedule);
  k.push_back((uintptr_t)h->opt_timing);
  for (int j = 0; j < kBuckets; ++j) k.push_back((uintptr_t)h->variant[j]);
  put(sa.xb), put(sa.xb32), put(sa.axb), put(sa.xa), put(sa.xa32), put(sa.axa);
  k.push_back((uintptr_t)sa.epi);
  putd(sa.polar_lat), putd(sa.gamma1), putd(sa.gamma2);
  put(sa.xa_out), put(sa.spread_b), put(sa.spread_a), put(sa.x_in);
  for (const DevBuf* b : {&h->cnt, &h->off, &h->scan_tmp, &h->row_idx, &h->row_gc, &h->row_wd,
                          &h->row_site, &h->nobsl_dev, &h->Abuf, &h->bbuf, &h->status, &h->trans,
                          &h->wout, &h->batch_tot, &h->b_order, &h->b_blk, &h->b_seg, &h->b_q,
                          &h->b_stats, &h->b_ticks, &h->eig.refl, &h->eig.tau, &h->eig.d,
                          &h->eig.e, &h->eig.lam, &h->eig.rot, &h->eig.hdr, &h->eig.nchase,
                          &h->eig.flags, &h->eig.rr_rot, &h->eig.rr_hdr, &h->eig.rr_nchase,
                          &h->eig.rr_lam})
    put(b->p);
  return k;
}

static void drop_graph(dida_letkf_handle* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->gkey.clear();
}

// Captures the call's launches (no timing events; the half-batch pipeline's
// second stream joins the capture through its fork / join events) unless
// the graph of an identical earlier call exists, and launches the graph.
static int graph_call(dida_letkf_handle* h, const StateArgs& sa, float& ms) {
  hipStream_t st = h->stream;
  if (!st) return set_error(DIDA_ERR_STATE, "graph mode needs a non-null stream");
  std::vector<uintptr_t> key = graph_key(h, sa);
  if (!h->gexec || key != h->gkey) {
    drop_graph(h);
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_call(h, sa, false, h->cfg.transform_kernel, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc < 0 || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc < 0) return rc;
      return set_error(DIDA_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    h->graph = g;
    HIP_TRY(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    h->gkey = key;
  }
  for (auto& e : h->g_ev)
    if (!e) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipEventRecord(h->g_ev[0], st));
  HIP_TRY(hipGraphLaunch(h->gexec, st));
  HIP_TRY(hipEventRecord(h->g_ev[1], st));
  HIP_TRY(hipEventSynchronize(h->g_ev[1]));
  HIP_TRY(hipEventElapsedTime(&ms, h->g_ev[0], h->g_ev[1]));
  return DIDA_OK;
}

static int run_analysis(dida_letkf_handle* h, const StateArgs& sa, dida_letkf_stats* stats) {
  if (!h->have_grid) return set_error(DIDA_ERR_STATE, "set_grid was not called");
  if (!h->have_obs) return set_error(DIDA_ERR_STATE, "set_obs was not called");
  if (h->cfg.overlap > 0 && (sa.xb32 || sa.epi))
    return set_error(DIDA_ERR_UNSUPPORTED,
                     "overlapped cubes update an FP64 state without the fused epilogue");
  RC_TRY(prepare_call(h));
  h->split_timed.assign(n_batches_of(h), false);
  h->run_capacity.assign(n_batches_of(h) * kBuckets, 0);
  bool graph = h->opt_graph != 0;
  bool reran = false;
  for (int attempt = 0;; ++attempt) {
    float gms = 0;
    const bool tune = !graph && h->opt_tune != 0 && h->opt_schedule == 1 && !h->tiled;
    if (graph) {
      RC_TRY(graph_call(h, sa, gms));
    } else {
      RC_TRY(enqueue_call(h, sa, true, h->cfg.transform_kernel, tune));
      if (tune) h->opt_tune = 0;
    }
    long long need = 0;
    const int flagged = finish_call(h, !graph, gms, sa.x_in != nullptr, stats, need);
    if (flagged < 0) return flagged;
    if (need <= h->row_cap) {
      if (stats) {
        stats->graph_replay = graph ? 1 : 0;
        stats->reran_rows = reran ? 1 : 0;
        stats->row_cap = h->row_cap;
      }
      return flagged;
    }
    // the lists were cut at the capacity: grow it to the measured need and
    // run the call again (eagerly; a graph is captured again on a later call)
    if (attempt > 0)
      return set_error(DIDA_ERR_STATE, "row lists grew between two passes over the same obs");
    h->row_cap = need + need / 8 + 4096;
    drop_graph(h);
    RC_TRY(prepare_call(h));
    graph = false;
    reran = true;
  }
}

extern "C" {

int dida_letkf_analyze_device(dida_letkf_handle* h, const double* xb, const double* axb,
                              double* xa, double* axa, dida_letkf_stats* stats) {
  if (!h || !xb || !axb || !xa || !axa) return set_error(DIDA_ERR_ARG, "null argument");
  StateArgs sa;
  sa.xb = xb;
  sa.axb = axb;
  sa.xa = xa;
  sa.axa = axa;
  return run_analysis(h, sa, stats);
}

int dida_letkf_analyze_device_f32(dida_letkf_handle* h, const float* xb, const double* axb,
                                  float* xa, double* axa, dida_letkf_stats* stats) {
  if (!h || !xb || !axb || !xa || !axa) return set_error(DIDA_ERR_ARG, "null argument");
  StateArgs sa;
  sa.xb32 = xb;
  sa.axb = axb;
  sa.xa32 = xa;
  sa.axa = axa;
  return run_analysis(h, sa, stats);
}

// analysis + fused epilogue; xb_f32 selects which of xb / xb32 is the state
static int analyze_epilogue(dida_letkf_handle* h, const double* xb, const float* xb32,
                            const double* axb, double polar_lat, double gamma1, double gamma2,
                            double* xa_out, double* spread_b, double* spread_a,
                            dida_letkf_stats* stats) {
  if (!h || (!xb && !xb32) || !axb || !xa_out || !spread_b || !spread_a)
    return set_error(DIDA_ERR_ARG, "null argument");
  const auto& c = h->cfg;
  StateArgs sa;
  sa.xb = xb;
  sa.xb32 = xb32;
  sa.axb = axb;
  sa.epi = 1;
  sa.polar_lat = polar_lat;
  sa.gamma1 = gamma1;
  sa.gamma2 = gamma2;
  sa.xa_out = xa_out;
  sa.spread_b = spread_b;
  sa.spread_a = spread_a;
  // the halo columns of xa_out are zero (module_da_parallel.F90:222-223)
  HIP_TRY(launch_zero_halo(h->stream, xa_out, c.nlat, c.nlon, c.halo, (long long)h->n * c.k));
  return run_analysis(h, sa, stats);
}

int dida_letkf_analyze_epilogue_device(dida_letkf_handle* h, const double* xb, const double* axb,
                                       double polar_lat, double gamma1, double gamma2,
                                       double* xa_out, double* spread_b, double* spread_a,
                                       dida_letkf_stats* stats) {
  if (!xb) return set_error(DIDA_ERR_ARG, "null argument");
  return analyze_epilogue(h, xb, nullptr, axb, polar_lat, gamma1, gamma2, xa_out, spread_b,
                          spread_a, stats);
}

int dida_letkf_analyze_epilogue_device_f32(dida_letkf_handle* h, const float* xb,
                                           const double* axb, double polar_lat, double gamma1,
                                           double gamma2, double* xa_out, double* spread_b,
                                           double* spread_a, dida_letkf_stats* stats) {
  if (!xb) return set_error(DIDA_ERR_ARG, "null argument");
  return analyze_epilogue(h, nullptr, xb, axb, polar_lat, gamma1, gamma2, xa_out, spread_b,
                          spread_a, stats);
}

int dida_letkf_split_device(dida_letkf_handle* h, const double* x_in, double* xb, float* xb32,
                            double* axb) {
  if (!h || !x_in || !axb || (!xb == !xb32))
    return set_error(DIDA_ERR_ARG, "null argument (give exactly one of xb, xb32)");
  const auto& c = h->cfg;
  hipStream_t st = h->stream;
  RC_TRY(h->sp_nan.ensure(2 * sizeof(unsigned long long)));
  SplitParams spp{};
  spp.x_in = x_in;
  spp.xb = xb;
  spp.xb32 = xb32;
  spp.axb = axb;
  spp.n_elems = (long long)(c.nlat + 2 * c.halo) * (c.nlon + 2 * c.halo) * h->n;
  spp.k = c.k;
  spp.nan = h->sp_nan.as<unsigned long long>();
  HIP_TRY(launch_split(st, spp));
  unsigned long long nan[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(nan, h->sp_nan.p, sizeof(nan), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (nan[0]) return nan_error(h, nan);
  return DIDA_OK;
}

int dida_letkf_run_da_device(dida_letkf_handle* h, const double* x_in, int state_f32,
                             double polar_lat, double gamma1, double gamma2, double* xa_out,
                             double* spread_b, double* spread_a, dida_letkf_stats* stats) {
  if (!h || !x_in || !xa_out || !spread_b || !spread_a)
    return set_error(DIDA_ERR_ARG, "null argument");
  const auto& c = h->cfg;
  const size_t ne = (size_t)(c.nlat + 2 * c.halo) * (c.nlon + 2 * c.halo) * h->n;
  RC_TRY(h->sp_nan.ensure(2 * sizeof(unsigned long long)));
  RC_TRY(h->sp_axb.ensure(ne * sizeof(double)));
  StateArgs sa;
  if (state_f32) {
    RC_TRY(h->sp_xb32.ensure(ne * c.k * sizeof(float)));
    sa.xb32 = h->sp_xb32.as<float>();
  } else {
    RC_TRY(h->sp_xb.ensure(ne * c.k * sizeof(double)));
    sa.xb = h->sp_xb.as<double>();
  }
  sa.axb = h->sp_axb.as<double>();
  sa.x_in = x_in;
  sa.epi = 1;
  sa.polar_lat = polar_lat;
  sa.gamma1 = gamma1;
  sa.gamma2 = gamma2;
  sa.xa_out = xa_out;
  sa.spread_b = spread_b;
  sa.spread_a = spread_a;
  HIP_TRY(launch_zero_halo(h->stream, xa_out, c.nlat, c.nlon, c.halo, (long long)h->n * c.k));
  return run_analysis(h, sa, stats);
}

int dida_letkf_analyze(dida_letkf_handle* h, const double* xb, const double* axb, double* xa,
                       double* axa, dida_letkf_stats* stats) {
  if (!h || !xb || !axb || !xa || !axa) return set_error(DIDA_ERR_ARG, "null argument");
  const auto& c = h->cfg;
  const size_t ncol = (size_t)(c.nlat + 2 * c.halo) * (c.nlon + 2 * c.halo) * h->n;
  const size_t nx = ncol * c.k;
  hipStream_t st = h->stream;
  RC_TRY(h->x_xb.ensure(nx * 8));
  RC_TRY(h->x_axb.ensure(ncol * 8));
  RC_TRY(h->x_xa.ensure(nx * 8));
  RC_TRY(h->x_axa.ensure(ncol * 8));
  HIP_TRY(hipMemcpyAsync(h->x_xb.p, xb, nx * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(h->x_axb.p, axb, ncol * 8, hipMemcpyHostToDevice, st));
  // outputs start as the caller's arrays so that halo entries are left as they are
  HIP_TRY(hipMemcpyAsync(h->x_xa.p, xa, nx * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(h->x_axa.p, axa, ncol * 8, hipMemcpyHostToDevice, st));
  StateArgs sa;
  sa.xb = h->x_xb.as<double>();
  sa.axb = h->x_axb.as<double>();
  sa.xa = h->x_xa.as<double>();
  sa.axa = h->x_axa.as<double>();
  int rc = run_analysis(h, sa, stats);
  if (rc < 0) return rc;
  HIP_TRY(hipMemcpyAsync(xa, h->x_xa.p, nx * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(axa, h->x_axa.p, ncol * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return rc;
}

int dida_letkf_get_transforms(dida_letkf_handle* h, long long n_an, double* trans,
                              double* overline_w, int* nobsl) {
  if (!h) return set_error(DIDA_ERR_ARG, "null handle");
  if (!h->cfg.keep_transforms) return set_error(DIDA_ERR_STATE, "keep_transforms is off");
  if (n_an != h->n_an) return set_error(DIDA_ERR_ARG, "n_analyses mismatch (%lld)", h->n_an);
  const int k = h->cfg.k;
  if (trans)
    HIP_TRY(hipMemcpyAsync(trans, h->trans.p, sizeof(double) * n_an * k * k, hipMemcpyDeviceToHost,
                           h->stream));
  if (overline_w)
    HIP_TRY(hipMemcpyAsync(overline_w, h->wout.p, sizeof(double) * n_an * k, hipMemcpyDeviceToHost,
                           h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (nobsl) std::memcpy(nobsl, h->nobsl_host.data(), sizeof(int) * n_an);
  return DIDA_OK;
}

}  // extern "C"

namespace {

// Device buffers of one batched letkf_core call: inputs staged from the host
// (Yb, rdiaginv, dep, rloc, the row list and offsets, the FP32 copy of Yb for
// the reduced modes), the products A and b, and the outputs.
struct CoreBufs {
  DevBuf y, r, d, g, idx, off, A, b, t, w, s, y32;
};

// Every device buffer the call needs with its size in bytes, in a fixed
// order: the core buffers, then the eigensolver's (as prepare_eig,
// eig_begin_call and rerun_flagged request them).
int core_layout(CoreBufs& cb, EigScratch& es, int k, long long n_an, long long nr, int precision,
                DevBuf** bufs, size_t* sizes) {
  const size_t kp = kp_of(k), n = n_an, nrr = std::max(nr, 1LL);
  const size_t cap = rerun_cap_of(k, n_an);
  int c = 0;
  auto add = [&](DevBuf& b, size_t bytes) { bufs[c] = &b; sizes[c] = bytes; ++c; };
  add(cb.y, 8 * nrr * k);
  add(cb.r, 4 * nrr);
  add(cb.d, 8 * nrr);
  add(cb.g, 8 * (nrr + kListPad));
  add(cb.idx, 4 * (nrr + kListPad));
  add(cb.off, 8 * (n + 1));
  add(cb.A, 8 * n * kp * kp);
  add(cb.b, 8 * n * kp);
  add(cb.t, 8 * n * k * k);
  add(cb.w, 8 * n * k);
  add(cb.s, 4 * n);
  if (precision != DIDA_PREC_FP64) add(cb.y32, 4 * (nrr + kListPad) * k);
  add(es.refl, 8 * n * kp * kp);
  add(es.tau, 8 * n * kp);
  add(es.d, 8 * n * kp);
  add(es.e, 8 * n * kp);
  add(es.lam, 8 * n * kp);
  add(es.rot, 16 * n * (size_t)rot_cap_of(k));
  add(es.hdr, 16 * n * (size_t)(kHdrCapPerK * k));
  add(es.nchase, 4 * n);
  add(es.flags, 4 * (2 + cap));
  add(es.rr_rot, 16 * cap * (size_t)(2 * rot_cap_of(k)));
  add(es.rr_hdr, 16 * cap * (size_t)(2 * kHdrCapPerK * k));
  add(es.rr_nchase, 4 * cap);
  add(es.rr_lam, 8 * cap * kp);
  return c;
}
constexpr int kCoreMaxBufs = 32;

int core_check(int k, long long n_an, const long long* row_ptr, const double* hxens,
               const float* rdiaginv, const double* dep, const double* rloc, double* trans,
               double* overline_w, double inflation, int precision) {
  if (k < 2 || k > 128) return set_error(DIDA_ERR_UNSUPPORTED, "k=%d: supported 2..128", k);
  if (n_an < 0 || !row_ptr || !trans || !overline_w)
    return set_error(DIDA_ERR_ARG, "bad arguments");
  if (precision < DIDA_PREC_FP64 || precision > DIDA_PREC_BF16X3)
    return set_error(DIDA_ERR_ARG, "bad precision");
  if (!(inflation > 0.0)) return set_error(DIDA_ERR_ARG, "inflation_factor must be > 0");
  if (n_an == 0) return DIDA_OK;
  if (row_ptr[0] != 0) return set_error(DIDA_ERR_ARG, "row_ptr[0] must be 0");
  for (long long a = 0; a < n_an; ++a)
    if (row_ptr[a + 1] < row_ptr[a]) return set_error(DIDA_ERR_ARG, "row_ptr decreasing");
  if (row_ptr[n_an] > 0 && (!hxens || !rdiaginv || !dep || !rloc))
    return set_error(DIDA_ERR_ARG, "null input arrays");
  return DIDA_OK;
}

// n_an independent letkf_core calls on st: stage the inputs, products,
// eigensolver (mode as run_transform), outputs back to the host.  The buffers
// of cb and es are sized here (a no-op for pieces they already hold).
int core_run(hipStream_t st, CoreBufs& cb, EigScratch& es, int mode, int k, long long n_an,
             const long long* row_ptr, const double* hxens, const float* rdiaginv,
             const double* dep, const double* rloc, double* trans, double* overline_w,
             double inflation, int precision, int* status) {
  const long long nr = row_ptr[n_an];
  const int kp = kp_of(k);
  if (precision != DIDA_PREC_FP64 && 4 * (nr + (long long)kListPad) * k >= kYbOffsetLimit)
    return set_error(DIDA_ERR_UNSUPPORTED, "%lld obs rows x k=%d exceed the 2 GiB of FP32 Yb "
                     "the reduced-precision products address", nr, k);
  DevBuf* bufs[kCoreMaxBufs];
  size_t sizes[kCoreMaxBufs];
  const int nb = core_layout(cb, es, k, n_an, nr, precision, bufs, sizes);
  // the core buffers; the eigensolver sizes its own
  for (int i = 0; i < nb; ++i)
    if (bufs[i] == &es.refl) break;
    else RC_TRY(bufs[i]->ensure(sizes[i]));
  if (nr > 0) {
    HIP_TRY(hipMemcpyAsync(cb.y.p, hxens, sizeof(double) * nr * k, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(cb.r.p, rdiaginv, sizeof(float) * nr, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(cb.d.p, dep, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(cb.g.p, rloc, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((nr + kListPad + 255) / 256)), dim3(256), 0, st,
                       nr, (long long)kListPad, cb.idx.as<int>());
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(cb.off.p, row_ptr, sizeof(long long) * (n_an + 1), hipMemcpyHostToDevice, st));
  if (precision != DIDA_PREC_FP64 && nr > 0) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((nr * k + 255) / 256)), dim3(256), 0, st,
                       nr * k, cb.y.as<double>(), cb.y32.as<float>());
    HIP_TRY(hipGetLastError());
  }
  Geom g{};
  g.nlat = 1; g.nlon = 1; g.cube_n = 1; g.cube_s = 1; g.nlat_c = 1; g.nlon_c = 1;
  ProdParams pp{};
  pp.g = g;
  pp.k = k;
  pp.kp = kp;
  pp.nzc = 1;
  pp.vertical = 0;
  pp.row_idx = cb.idx.as<int>();
  pp.row_gc = cb.g.as<double>();
  pp.row_off = cb.off.as<long long>();
  pp.row_cap = nr;
  pp.groups = 1;
  pp.yb64 = cb.y.as<double>();
  pp.yb32 = precision != DIDA_PREC_FP64 ? cb.y32.as<float>() : nullptr;
  pp.yb32_bytes = (unsigned)std::min<size_t>(cb.y32.bytes, kYbOffsetLimit);
  pp.dep = cb.d.as<double>();
  pp.rinv32 = cb.r.as<float>();
  pp.a0 = 0;
  pp.n_an = n_an;
  pp.h0 = 0;
  pp.A = cb.A.as<double>();
  pp.b = cb.b.as<double>();
  HIP_TRY(launch_products(st, pp, precision));
  XformParams xp{};
  xp.g = g;
  xp.k = k;
  xp.kp = kp;
  xp.a_f32 = precision == DIDA_PREC_BF16X3 && kp <= 64;
  xp.inflation = inflation;
  xp.guard = precision != DIDA_PREC_FP64;
  xp.A = pp.A;
  xp.b = pp.b;
  xp.a0 = 0;
  xp.n_an = n_an;
  xp.nzc = 1;
  xp.xb = nullptr;
  xp.trans_out = cb.t.as<double>();
  xp.w_out = cb.w.as<double>();
  xp.status = cb.s.as<int>();
  xp.list = nullptr;
  RC_TRY(eig_begin_call(st, es, k, n_an));
  RC_TRY(run_transform(st, es, xp, mode));
  HIP_TRY(hipMemcpyAsync(trans, cb.t.p, sizeof(double) * n_an * k * k, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(overline_w, cb.w.p, sizeof(double) * n_an * k, hipMemcpyDeviceToHost, st));
  // status in chunks through a stack buffer when the caller passes none
  int flagged = 0;
  int chunk[1024];
  for (long long a0 = 0; a0 < n_an; a0 += 1024) {
    const long long c = std::min<long long>(1024, n_an - a0);
    int* dst = status ? status + a0 : chunk;
    HIP_TRY(hipMemcpyAsync(dst, cb.s.as<int>() + a0, sizeof(int) * c, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (long long i = 0; i < c; ++i) flagged += dst[i] != 0;
  }
  return flagged;
}

void release_core(CoreBufs& cb, EigScratch& es) {
  DevBuf* bufs[kCoreMaxBufs];
  size_t sizes[kCoreMaxBufs];
  const int nb = core_layout(cb, es, 2, 1, 1, DIDA_PREC_BF16X3, bufs, sizes);
  for (int i = 0; i < nb; ++i) bufs[i]->release();
}

// The drop-in letkf_core_ is called once per local analysis (64,800 times a
// cycle at 360x180): each calling host thread keeps, per device, a stream and
// the buffers of its largest call so far, until the process ends, so that a
// call costs its copies and launches only.
struct CoreCtx {
  int device = -1;
  hipStream_t st = nullptr;
  CoreBufs cb;
  EigScratch es;
};

int core_ctx(CoreCtx** out) {
  thread_local std::vector<CoreCtx*> ctxs;   // never freed: see above
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  for (CoreCtx* c : ctxs)
    if (c->device == dev) {
      *out = c;
      return DIDA_OK;
    }
  CoreCtx* c = new (std::nothrow) CoreCtx();
  if (!c) return set_error(DIDA_ERR_NOMEM, "host allocation failed");
  c->device = dev;
  const hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_error(DIDA_ERR_HIP, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
  }
  ctxs.push_back(c);
  *out = c;
  return DIDA_OK;
}

}  // namespace

extern "C" {

int dida_letkf_core_batched(int k, long long n_an, const long long* row_ptr, const double* hxens,
                            const float* rdiaginv, const double* dep, const double* rloc,
                            double* trans, double* overline_w, double inflation, int precision,
                            int* status) {
  RC_TRY(core_check(k, n_an, row_ptr, hxens, rdiaginv, dep, rloc, trans, overline_w, inflation,
                    precision));
  if (n_an == 0) return DIDA_OK;
  hipStream_t st;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  CoreBufs cb;
  EigScratch es;
  const int rc = core_run(st, cb, es, 0, k, n_an, row_ptr, hxens, rdiaginv, dep, rloc, trans,
                          overline_w, inflation, precision, status);
  (void)hipStreamSynchronize(st);
  es.release_streams();
  release_core(cb, es);
  (void)hipStreamDestroy(st);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// dida_letkf_core_batched_noalloc: dida_letkf_core_batched with every device
// buffer carved from `scratch`, a fixed array in the library's code object, so
// the call itself allocates no device memory.
// ---------------------------------------------------------------------------

// 64 MiB on each device that runs the library, held until the process ends.
// In host code this name is only a handle for hipGetSymbolAddress: the
// device's copy is reached through the pointer that call returns.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
struct NoAllocDevice {
  std::mutex mu;            // calls on one device run one after another
  void* ptr = nullptr;      // the device's copy of scratch
};
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

constexpr size_t kNoAllocAlign = 256;
size_t round_up(size_t n) { return (n + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }

// Scratch bytes a call needs: every buffer of core_layout, 256-byte aligned.
size_t noalloc_bytes(int k, long long n_an, long long nr, int precision) {
  CoreBufs cb;
  EigScratch es;
  DevBuf* bufs[kCoreMaxBufs];
  size_t sizes[kCoreMaxBufs];
  const int nb = core_layout(cb, es, k, n_an, nr, precision, bufs, sizes);
  size_t total = 0;
  for (int i = 0; i < nb; ++i) total += round_up(sizes[i]);
  return total;
}

}  // namespace

extern "C" {

int dida_letkf_core_batched_noalloc(int k, long long n_an, const long long* row_ptr,
                                    const double* hxens, const float* rdiaginv, const double* dep,
                                    const double* rloc, double* trans, double* overline_w,
                                    double inflation, int precision, int* status) {
  RC_TRY(core_check(k, n_an, row_ptr, hxens, rdiaginv, dep, rloc, trans, overline_w, inflation,
                    precision));
  if (n_an == 0) return DIDA_OK;
  const size_t need = noalloc_bytes(k, n_an, row_ptr[n_an], precision);
  if (need > kNoAllocScratchBytes)
    return set_error(DIDA_ERR_ARG,
                     "dida_letkf_core_batched_noalloc: %lld analyses at k=%d with %lld obs rows "
                     "need %zu bytes of scratch, more than its fixed %zu",
                     n_an, k, row_ptr[n_an], need, kNoAllocScratchBytes);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return set_error(DIDA_ERR_UNSUPPORTED, "device %d: the no-alloc path covers devices 0..%d", dev,
                     kNoAllocMaxDevices - 1);
  NoAllocDevice& D = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(D.mu);
  if (!D.ptr) {
    // set-up, once per device with the device current
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(DIDA_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null pointer");
    D.ptr = p;
  }
  CoreBufs cb;
  EigScratch es;
  DevBuf* bufs[kCoreMaxBufs];
  size_t sizes[kCoreMaxBufs];
  const int nb = core_layout(cb, es, k, n_an, row_ptr[n_an], precision, bufs, sizes);
  unsigned char* base = static_cast<unsigned char*>(D.ptr);
  size_t off = 0;
  for (int i = 0; i < nb; ++i) {
    bufs[i]->wrap(base + off, sizes[i]);
    off += round_up(sizes[i]);
  }
  // one stream, the null stream: no stream or event is created
  const int rc = core_run(nullptr, cb, es, 2, k, n_an, row_ptr, hxens, rdiaginv, dep, rloc,
                          trans, overline_w, inflation, precision, status);
  (void)hipStreamSynchronize(nullptr);
  release_core(cb, es);       // forgets the wrapped pieces; frees nothing
  return rc;
}

void letkf_core_(const int* nanals, const int* nobsl, const double* hxens_in,
                 const float* rdiaginv, const double* dep, const double* rloc, double* trans,
                 double* overline_w, const double* inflation_factor, const int* pflag) {
  (void)pflag;
  long long rp[2] = {0, (long long)*nobsl};
  int st = 0;
  CoreCtx* cx = nullptr;
  int rc = core_check(*nanals, 1, rp, hxens_in, rdiaginv, dep, rloc, trans, overline_w,
                      *inflation_factor, DIDA_PREC_FP64);
  if (rc == DIDA_OK) rc = core_ctx(&cx);
  if (rc == DIDA_OK) {
    rc = core_run(cx->st, cx->cb, cx->es, 0, *nanals, 1, rp, hxens_in, rdiaginv, dep, rloc,
                  trans, overline_w, *inflation_factor, DIDA_PREC_FP64, &st);
    const hipError_t e = hipStreamSynchronize(cx->st);
    if (rc >= 0 && e != hipSuccess)
      rc = set_error(DIDA_ERR_HIP, "letkf_core_: %s", hipGetErrorString(e));
  }
  if (rc < 0 && *nanals > 0) {
    // no result: the outputs are NaN, so the caller's update cannot pass for one
    const double nan = std::nan("");
    for (long long i = 0; i < (long long)*nanals * *nanals; ++i) trans[i] = nan;
    for (int i = 0; i < *nanals; ++i) overline_w[i] = nan;
  }
  if (rc != 0) fprintf(stderr, "letkf_core_: rc=%d %s\n", rc, g_last_error.c_str());
}

size_t dida_letkf_job_bytes(const dida_letkf_config* c, int n_sites, int n_rows) {
  if (!c || c->k < 2 || c->nlat < 1 || c->nlon < 1 || c->cube_n < 1) return 0;
  const size_t k = c->k, kp = kp_of(c->k);
  const size_t ns = std::max(n_sites, 1), nr = std::max(n_rows, 1);
  const int cz = c->cube_nz > 0 ? c->cube_nz : c->nz;
  const size_t nzc = c->vert_dist > 0.0 ? (c->nz + cz - 1) / cz : 1;
  const size_t mb = (size_t)max_batch_columns(*c);
  const size_t nh = (size_t)cubes_h(*c);
  const size_t n_an = nh * nzc;
  const size_t ovl_trans = c->overlap > 0 ? n_an * k * k * 8 : 0;
  const size_t n_state = (size_t)(c->nlat + 2 * c->halo) * (c->nlon + 2 * c->halo) *
                         (c->n2d + c->n3d * c->nz);
  size_t b = 0;
  b += ns * (8 * 2 + 4 + 16 + 24 + 8) + 16 * ns + (1 << 20);       // sites, keys, index, sort
  b += nr * (8 * k + 8 * 3 + 4);                                    // rows
  if (c->precision != DIDA_PREC_FP64) b += nr * 4 * k;              // FP32 copy of yb
  b += 2 * 4 * 200000;                                              // bins (upper bound)
  // local obs rows of one batch (row_idx, row_gc, row_wd), provisioned at
  // six times the uniform density: a call whose lists need more runs again
  // with the measured need, within this
  b += 20 * mb * (size_t)rows_per_column_bound(*c, n_rows);
  b += 16 * (mb + 1) * 2;
  b += mb * nzc * (8 * kp * kp + 8 * kp + eig_bytes_per_analysis(c->k));
  b += rerun_cap_of(c->k, (long long)(mb * nzc)) * rerun_bytes_per_slot(c->k);
  b += 4 * n_an;
  if (c->keep_transforms) b += n_an * (8 * k * k + 8 * k);
  else b += ovl_trans;                                              // overlapped cubes' trans
  b += n_state * (k + 1) * 8;                                       // serving ping-pong state
  b += nr * (k + 1) * 8;                                            // serving operator Yb, d
  return b + (16 << 20);
}

int dida_letkf_fill_halo_device(dida_letkf_handle* h, double* x, double* ax) {
  if (!h) return set_error(DIDA_ERR_ARG, "null handle");
  const auto& c = h->cfg;
  if (x) HIP_TRY(launch_fill_halo(h->stream, x, c.nlat, c.nlon, c.halo, (long long)h->n * c.k));
  if (ax) HIP_TRY(launch_fill_halo(h->stream, ax, c.nlat, c.nlon, c.halo, (long long)h->n));
  return DIDA_OK;
}

int dida_letkf_obs_operator_device(dida_letkf_handle* h, const double* xb, const double* axb,
                                   int n_sites, const double* site_lat, const double* site_lon,
                                   const int* site_row_ptr, int n_rows, const int* row_var,
                                   const double* row_ph, const double* yo, double* yb,
                                   double* dep) {
  if (!h) return set_error(DIDA_ERR_ARG, "null handle");
  if (n_sites < 0 || n_rows < 0) return set_error(DIDA_ERR_ARG, "negative sizes");
  if (n_sites > 0 && (!xb || !axb || !site_lat || !site_lon || !site_row_ptr))
    return set_error(DIDA_ERR_ARG, "null argument");
  if (n_rows > 0 && (!row_var || !row_ph || !yo || !yb || !dep))
    return set_error(DIDA_ERR_ARG, "null argument");
  if (!h->have_grid) return set_error(DIDA_ERR_STATE, "set_grid was not called");
  const auto& c = h->cfg;
  if (c.halo < 1) return set_error(DIDA_ERR_ARG, "the observation operator needs halo >= 1");
  if (h->hyai.empty()) return set_error(DIDA_ERR_ARG, "the observation operator needs hyai/hybi");
  if (c.ps_index < 0 || c.ps_index >= h->n) return set_error(DIDA_ERR_ARG, "bad ps_index");
  ObsOpParams p{};
  p.k = c.k;
  p.n = h->n;
  p.nz = c.nz;
  p.nlat = c.nlat;
  p.nlon = c.nlon;
  p.halo = c.halo;
  p.ps_index = c.ps_index;
  p.n_var = c.n2d + c.n3d;
  p.n_rows = n_rows;
  int cur = 1;
  for (int v = 0; v < c.n2d; ++v) { p.var_start[v] = cur; p.var_lev[v] = 1; cur += 1; }
  for (int v = 0; v < c.n3d; ++v) {
    p.var_start[c.n2d + v] = cur;
    p.var_lev[c.n2d + v] = c.nz;
    cur += c.nz;
  }
  p.glat = h->glat.as<float>();
  p.glon = h->glon.as<float>();
  p.hyai = h->d_hyai.as<double>();
  p.hybi = h->d_hybi.as<double>();
  p.p0 = c.p0;
  p.ptop = c.ptop;
  p.xb = xb;
  p.axb = axb;
  p.n_sites = n_sites;
  p.site_lat = site_lat;
  p.site_lon = site_lon;
  p.site_row_ptr = site_row_ptr;
  p.row_var = row_var;
  p.row_ph = row_ph;
  p.yo = yo;
  p.yb = yb;
  p.dep = dep;
  HIP_TRY(launch_obs_operator(h->stream, p));
  return DIDA_OK;
}

int dida_letkf_epilogue_device(dida_letkf_handle* h, const double* xb, const double* axb,
                               const double* xa, const double* axa, double polar_lat,
                               double gamma1, double gamma2, double* xa_out, double* spread_b,
                               double* spread_a) {
  if (!h || !xb || !axb || !xa || !axa || !xa_out || !spread_b || !spread_a)
    return set_error(DIDA_ERR_ARG, "null argument");
  if (!h->have_grid) return set_error(DIDA_ERR_STATE, "set_grid was not called");
  EpiParams p;
  p.k = h->cfg.k;
  p.n = h->n;
  p.nlat = h->cfg.nlat;
  p.nlon = h->cfg.nlon;
  p.halo = h->cfg.halo;
  p.glat = h->glat.as<float>();
  p.xb = xb;
  p.axb = axb;
  p.xa = xa;
  p.axa = axa;
  p.polar_lat = polar_lat;
  p.gamma1 = gamma1;
  p.gamma2 = gamma2;
  p.xa_out = xa_out;
  p.spread_b = spread_b;
  p.spread_a = spread_a;
  HIP_TRY(launch_epilogue(h->stream, p));
  return DIDA_OK;
}

}  // extern "C"
