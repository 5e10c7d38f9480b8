// letkf_obsop.hip -- observation operator H and Yb/d centring on device.
//
// One wavefront per obs site, lane e = member e (members e and e + 64 for
// k > 64).  Per site: the grid cell (find_obs_grid, letkf_obs.F90:192-248),
// each member's pressure profile from its interpolated ps
// (ps2ph_3d_interp, obs_make.F90:345-381), then per row the bilinear value,
// with linear interpolation in ln p between the bracketing levels for 3-D
// variables (obs_interpolation_3d_2, obs_make.F90:233-342); the level search
// continues from the previous row of the same site, as the reference's does
// across the levels of one obs.  Finally Yb = H(x) - mean, d = yo - mean
// (letkf_obs.F90:158-170).
#include <algorithm>

#include "dida_letkf.h"

#include "letkf_device.h"
#include "letkf_kernels.h"

namespace dida {

// first 1-based index i in [1, n + 1] with grid(i) >= v, grid(n + 1) = close;
// n + 1 also for v above close (outside the grid: clamped to the closing cell)
__device__ __forceinline__ int grid_index(const float* __restrict__ g, int n, float close, double v) {
  int lo = 0, hi = n;                  // answer in [lo, hi] as a 0-based count of entries < v
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((double)g[mid] >= v) hi = mid;
    else lo = mid + 1;
  }
  (void)close;
  return lo + 1;
}

__device__ __forceinline__ double grid_weight(const float* __restrict__ g, int n, float close, int t,
                                              double v) {
#pragma clang fp contract(off)
  if (t <= 1) return 0.0;
  const float ga = g[t - 2];
  const float gb = (t <= n) ? g[t - 1] : close;
  const float df = gb - ga;            // real(4) spacing (letkf_obs.F90:234)
  const double d = (double)df;
  return (d - v + (double)ga) / d;
}

__device__ __forceinline__ double ph_mid(const ObsOpParams& p, int l, double ps) {
#pragma clang fp contract(off)
  // mid-level l (1-based) of the hybrid coordinate (vert_coord_mod.F90:556-561);
  // level nz + 1 repeats level nz
  if (l > p.nz) l = p.nz;
  const double a = p.hyai[l - 1] * (p.p0 - p.ptop) + p.hybi[l - 1] * (ps - p.ptop) + p.ptop;
  const double b = p.hyai[l] * (p.p0 - p.ptop) + p.hybi[l] * (ps - p.ptop) + p.ptop;
  return 0.5 * (a + b);
}

template <int MEMB>
__global__ __launch_bounds__(256) void k_obs_operator(ObsOpParams p) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= p.n_sites) return;
  const int k = p.k, n = p.n, h = p.halo;
  int tx = grid_index(p.glat, p.nlat, 90.0f, p.site_lat[s]);
  int ty = grid_index(p.glon, p.nlon, 360.0f, p.site_lon[s]);
  const double ax = grid_weight(p.glat, p.nlat, 90.0f, tx, p.site_lat[s]);
  const double ay = grid_weight(p.glon, p.nlon, 360.0f, ty, p.site_lon[s]);
  const long long lat_ext = p.nlat + 2 * h;
  // state offsets of the points (tx-1 | tx) x (ty-1 | ty)
  const long long o00 = ((long long)(ty - 2 + h) * lat_ext + (tx - 2 + h)) * n;
  const long long o10 = ((long long)(ty - 2 + h) * lat_ext + (tx - 1 + h)) * n;
  const long long o01 = ((long long)(ty - 1 + h) * lat_ext + (tx - 2 + h)) * n;
  const long long o11 = ((long long)(ty - 1 + h) * lat_ext + (tx - 1 + h)) * n;
  auto bil = [&](int el, int e) -> double {
    const double x00 = p.xb[(o00 + el) * k + e] + p.axb[o00 + el];
    const double x10 = p.xb[(o10 + el) * k + e] + p.axb[o10 + el];
    const double x01 = p.xb[(o01 + el) * k + e] + p.axb[o01 + el];
    const double x11 = p.xb[(o11 + el) * k + e] + p.axb[o11 + el];
    const double r1 = x00 * ax + x10 * (1 - ax);
    const double r2 = x01 * ax + x11 * (1 - ax);
    return r1 * ay + r2 * (1 - ay);
  };
  double ps[MEMB];
  int jp[MEMB];
#pragma unroll
  for (int m = 0; m < MEMB; ++m) {
    const int e = lane + 64 * m;
    ps[m] = (e < k) ? bil(p.ps_index, e) : 1.0e5;
    jp[m] = 1;
  }
  // rows outside [0, n_rows) are skipped; a row with an unknown variable
  // gets NaN (bad input never reads outside the state arrays)
  const int r0 = max(p.site_row_ptr[s], 0), r1 = min(p.site_row_ptr[s + 1], p.n_rows);
  for (int r = r0; r < r1; ++r) {
    const int v = p.row_var[r];
    if (v < 0 || v >= p.n_var) {
#pragma unroll
      for (int m = 0; m < MEMB; ++m) {
        const int e = lane + 64 * m;
        if (e < k) p.yb[(long long)r * k + e] = __builtin_nan("");
      }
      if (lane == 0) p.dep[r] = __builtin_nan("");
      continue;
    }
    const int vs = p.var_start[v] - 1;
    const bool three_d = p.var_lev[v] > 1;
    const double pr = p.row_ph[r];
    double val[MEMB];
    double sum = 0.0;
#pragma unroll
    for (int m = 0; m < MEMB; ++m) {
      const int e = lane + 64 * m;
      double x = 0.0;
      if (e < k) {
        int j = jp[m];
        while (j <= p.nz && pr > ph_mid(p, j, ps[m])) ++j;
        jp[m] = j;
        const int pos = j - 1;
        if (three_d) {
          if (pos >= 1 && pos < p.nz) {
            const int tl = vs + pos - 1;
            const double lu = log(ph_mid(p, pos + 1, ps[m])), ll = log(ph_mid(p, pos, ps[m]));
            const double t3 = (lu - log(pr)) / (lu - ll);
            x = bil(tl, e) * t3 + bil(tl + 1, e) * (1 - t3);
          } else {
            x = bil(pos == 0 ? vs : vs + p.nz - 1, e);
          }
        } else {
          x = bil(vs, e);
        }
      }
      val[m] = x;
      sum += x;
    }
    const double mean = wave_sum(sum) / k;
#pragma unroll
    for (int m = 0; m < MEMB; ++m) {
      const int e = lane + 64 * m;
      if (e < k) p.yb[(long long)r * k + e] = val[m] - mean;
    }
    if (lane == 0) p.dep[r] = p.yo[r] - mean;
  }
}

// One workgroup per halo column: periodic in longitude, the nearest interior
// row past either pole; each copies `inner` contiguous values.
__global__ void k_fill_halo(double* __restrict__ x, int nlat, int nlon, int halo,
                            long long inner) {
  const int lat_ext = nlat + 2 * halo, lon_ext = nlon + 2 * halo;
  const int id = blockIdx.x;
  if (id >= lat_ext * lon_ext) return;
  const int lo = id / lat_ext, la = id % lat_ext;
  const bool in_lo = lo >= halo && lo < nlon + halo, in_la = la >= halo && la < nlat + halo;
  if (in_lo && in_la) return;
  int ls = lo;
  if (lo < halo) ls = lo + nlon;
  else if (lo >= nlon + halo) ls = lo - nlon;
  const int as = min(max(la, halo), nlat + halo - 1);
  const double* src = x + ((long long)ls * lat_ext + as) * inner;
  double* dst = x + ((long long)lo * lat_ext + la) * inner;
  for (long long i = threadIdx.x; i < inner; i += blockDim.x) dst[i] = src[i];
}

// Zeroes the halo columns of a [lon + 2h][lat + 2h][inner] field.
__global__ void k_zero_halo(double* __restrict__ x, int nlat, int nlon, int halo, long long inner) {
  const long long lat_ext = nlat + 2 * halo, lon_ext = nlon + 2 * halo;
  const long long total = lat_ext * lon_ext * inner;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long col = i / inner;
    const long long la = col % lat_ext, lo = col / lat_ext;
    const bool interior = la >= halo && la < nlat + halo && lo >= halo && lo < nlon + halo;
    if (!interior) x[i] = 0.0;
  }
}

hipError_t launch_zero_halo(hipStream_t st, double* x, int nlat, int nlon, int halo, long long inner) {
  if (halo <= 0) return hipSuccess;
  const long long total = (long long)(nlat + 2 * halo) * (nlon + 2 * halo) * inner;
  const unsigned nb = (unsigned)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_zero_halo, dim3(nb), dim3(256), 0, st, x, nlat, nlon, halo, inner);
  return hipGetLastError();
}

hipError_t launch_fill_halo(hipStream_t st, double* x, int nlat, int nlon, int halo, long long inner) {
  if (halo <= 0 || inner <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((nlat + 2 * halo) * (nlon + 2 * halo));
  hipLaunchKernelGGL(k_fill_halo, dim3(blocks), dim3(256), 0, st, x, nlat, nlon, halo, inner);
  return hipGetLastError();
}

hipError_t launch_obs_operator(hipStream_t st, const ObsOpParams& p) {
  if (p.n_sites <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((p.n_sites + 3) / 4);
  if (p.k <= 64) hipLaunchKernelGGL(k_obs_operator<1>, dim3(blocks), dim3(256), 0, st, p);
  else if (p.k <= 128) hipLaunchKernelGGL(k_obs_operator<2>, dim3(blocks), dim3(256), 0, st, p);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace dida

// ---------------------------------------------------------------------------
// Observation records as read_obs receives them (read_obs_mod.F90:338-449):
// host-side conversion into the sites and rows of the observation operator.
// ---------------------------------------------------------------------------

namespace {

// record i starts a new site: every record of a single-level type; on a
// multi-level type (ADPUPA, GPSRO) the records with lev == 1 (:411-416)
bool starts_site(int multi_level, float lev) { return !multi_level || lev == 1.0f; }

}  // namespace

extern "C" {

int dida_obs_records_count(int n_records, int nvar, int multi_level, const float* lev,
                           int* n_sites, int* n_rows) {
  if (n_records < 0 || nvar < 1 || !n_sites || !n_rows || (n_records > 0 && multi_level && !lev))
    return dida::set_error(DIDA_ERR_ARG, "obs records: bad arguments");
  int s = 0;
  for (int i = 0; i < n_records; ++i) {
    const float l = multi_level ? lev[i] : 1.0f;
    if (starts_site(multi_level, l)) ++s;
    else if (s == 0)
      return dida::set_error(DIDA_ERR_ARG, "obs records: record %d continues a profile that "
                             "never started (lev = %g before any lev = 1)", i, (double)l);
    if (!(l >= 1.0f) || l != (float)(int)l)
      return dida::set_error(DIDA_ERR_ARG, "obs records: record %d has level %g", i, (double)l);
  }
  *n_sites = s;
  *n_rows = n_records * nvar;
  return DIDA_OK;
}

int dida_obs_records_convert(int n_records, int nvar, int multi_level, const float* lon,
                             const float* lat, const float* lev, const float* val,
                             const float* noise, const int* var_index, double* site_lat,
                             double* site_lon, int* site_row_ptr, int* row_var, double* row_ph,
                             double* yo, double* rinv) {
  int ns = 0, nr = 0;
  const int rc = dida_obs_records_count(n_records, nvar, multi_level, lev, &ns, &nr);
  if (rc < 0) return rc;
  if (n_records > 0 && (!lon || !lat || !val || !var_index || !site_lat || !site_lon ||
                        !site_row_ptr || !row_var || !row_ph || !yo || !rinv))
    return dida::set_error(DIDA_ERR_ARG, "obs records: null argument");
  const int w = 2 * nvar + 1;     // val_buf(2 nvar + 1, obs_num): value, oe per var, then ph
  int s = -1;
  for (int i = 0; i < n_records; ++i) {
    if (starts_site(multi_level, multi_level ? lev[i] : 1.0f)) {
      ++s;
      site_lat[s] = lat[i];
      site_lon[s] = lon[i];
      site_row_ptr[s] = i * nvar;
    }
    const float* v = val + (size_t)i * w;
    for (int n = 0; n < nvar; ++n) {
      const int r = i * nvar + n;
      // :400-405 value + randn * oe, and :430-442 rinv = 1 / oe**2, in real(4)
      const float oe = v[2 * n + 1];
      const float z = noise ? noise[(size_t)i * nvar + n] : 0.0f;
      yo[r] = (double)(v[2 * n] + z * oe);
      rinv[r] = (double)(1.0f / (oe * oe));
      row_var[r] = var_index[n];
      row_ph[r] = (double)v[2 * nvar];
    }
  }
  if (ns > 0) site_row_ptr[ns] = n_records * nvar;
  return DIDA_OK;
}

}  // extern "C"
