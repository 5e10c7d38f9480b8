// letkf_kernels.hip -- gfx950 kernels of the LETKF analysis step.
//
//   select      local obs selection (gen_yb_dis, letkf_yb.F90:3-150) over a
//               lat-lon bin index of the obs sites, one wavefront per column
//   products    obs-space products A = Yb^T R^-1 rho Yb and
//               b = Yb^T R^-1 rho d (letkf_core.F90:58-79,146-195) on MFMA
//               (f64 16x16x4, f32 16x16x4, bf16 16x16x32 hi/lo split)
//   transform   FP64 eigensolve of A, Pa, T = sqrt((k-1) Pa), overline_w and
//               trans = T + overline_w 1^T (letkf_core.F90:97-225), fused with
//               the ensemble update xb_2_xa (letkf_compute.F90:259-329) or the
//               per-level update of the vertical loop (letkf_compute.F90:501-552)
//
// MFMA result layouts on gfx950 (checked on the device):
//   f64 16x16x4 : lane l, reg r holds D[l/16 + 4r][l%16]
//   f32 16x16x4, bf16 16x16x32 : lane l, reg r holds D[4(l/16) + r][l%16]
// Operands: A[i][k] at lane i + 16*(k/K0), B[k][j] at lane j + 16*(k/K0),
// with K0 = 1 (x4 shapes) or 8 consecutive k per lane (16x16x32 bf16).
#include "letkf_kernels.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <type_traits>
#include "letkf_device.h"

namespace dida {

// ---------------------------------------------------------------------------
// obs site index
// ---------------------------------------------------------------------------

__global__ void k_site_keys(int n_sites, const double* __restrict__ lat,
                            const double* __restrict__ lon, double obs_polar, int nb_lat,
                            int nb_lon, double bdeg_lat, double bdeg_lon, int* __restrict__ key,
                            int* __restrict__ val, int* __restrict__ hist) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sites) return;
  double la = lat[s];
  int kk = nb_lat * nb_lon;  // sentinel: dropped (letkf_yb.F90:87)
  if (fabs(la) <= obs_polar) {
    double lo = fmod(lon[s], 360.0);
    if (lo < 0.0) lo += 360.0;
    int b = (int)floor((la + 90.0) / bdeg_lat);
    b = b < 0 ? 0 : (b >= nb_lat ? nb_lat - 1 : b);
    int j = (int)floor(lo / bdeg_lon);
    j = j < 0 ? 0 : (j >= nb_lon ? nb_lon - 1 : j);
    kk = b * nb_lon + j;
    atomicAdd(&hist[kk], 1);
  }
  key[s] = kk;
  val[s] = s;
}

__global__ void k_site_gather(int n_sites, const int* __restrict__ order,
                              const double* __restrict__ lat, const double* __restrict__ lon,
                              const int* __restrict__ row_ptr, double* __restrict__ s_cos,
                              double* __restrict__ s_sin, double* __restrict__ s_lonr,
                              double* __restrict__ s_cosl, double* __restrict__ s_sinl,
                              int* __restrict__ s_row0, int* __restrict__ s_nrow) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_sites) return;
  int s = order[i];
  double xr = deg2rad_ref(lat[s]);
  s_cos[i] = cos(xr);
  s_sin[i] = sin(xr);
  const double lr = deg2rad_ref(lon[s]);
  s_lonr[i] = lr;
  s_cosl[i] = cos(lr);
  s_sinl[i] = sin(lr);
  int r0 = row_ptr ? row_ptr[s] : s;
  int r1 = row_ptr ? row_ptr[s + 1] : s + 1;
  s_row0[i] = r0;
  s_nrow[i] = r1 - r0;
}

// ---------------------------------------------------------------------------
// local obs selection
// ---------------------------------------------------------------------------

__device__ __forceinline__ void column_center(const Geom& g, int ah, int* sj, int* ej, int* si,
                                              int* ei) {
  const int ic = ah % g.nlat_c, jc = ah / g.nlat_c;
  cube_span(g, ic, g.nlat, sj, ej);
  cube_span(g, jc, g.nlon, si, ei);
}

// Bins a disc of half-angle th (degrees) around (clat, clon) touches: latitude
// bands [b0, b1] and one or two longitude bin ranges [ja[r], jb[r]].
struct BinRange {
  int b0, b1, nr;
  int ja[2], jb[2];
};

__device__ __forceinline__ BinRange bin_range(const SelectParams& p, double clat, double clon,
                                              double th) {
  BinRange R;
  R.b0 = (int)floor((clat - th + 90.0) / p.bdeg_lat);
  R.b1 = (int)floor((clat + th + 90.0) / p.bdeg_lat);
  R.b0 = R.b0 < 0 ? 0 : R.b0;
  R.b1 = R.b1 >= p.nb_lat ? p.nb_lat - 1 : R.b1;
  bool full = fabs(clat) + th >= 89.99;
  double dlam = 360.0;
  if (!full) {
    double sr = sin(th * (M_PI / 180.0)) / cos(fabs(clat) * (M_PI / 180.0));
    if (sr >= 1.0) full = true;
    else dlam = asin(sr) * (180.0 / M_PI) + 1e-3;
  }
  double clw = fmod(clon, 360.0);
  if (clw < 0.0) clw += 360.0;
  R.nr = 1;
  R.ja[0] = 0;
  R.jb[0] = p.nb_lon - 1;
  if (!full) {
    int j0 = (int)floor((clw - dlam) / p.bdeg_lon);
    int j1 = (int)floor((clw + dlam) / p.bdeg_lon);
    if (j1 - j0 + 1 < p.nb_lon) {
      if (j0 < 0) {
        R.ja[0] = j0 + p.nb_lon; R.jb[0] = p.nb_lon - 1; R.ja[1] = 0; R.jb[1] = j1; R.nr = 2;
      } else if (j1 >= p.nb_lon) {
        R.ja[0] = j0; R.jb[0] = p.nb_lon - 1; R.ja[1] = 0; R.jb[1] = j1 - p.nb_lon; R.nr = 2;
      } else {
        R.ja[0] = j0; R.jb[0] = j1;
      }
    }
  }
  return R;
}

// orders a wavefront's LDS writes before its reads of other lanes' entries
// (the wavefronts of k_select's workgroup never share LDS entries)
__device__ __forceinline__ void __syncwarp_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A candidate site is inside the support iff its exact latlon_dist_r8 gives
// ri <= 2 (schr and schr2 are positive up to ri = 2 and zero beyond).  In
// v = cos(angle) that is v >= cos(2c/a) up to the rounding of the exact
// evaluation, so a cheap v (cos of the longitude difference expanded with
// the sites' precomputed cos/sin of longitude, a few ulps from the exact v)
// decides every candidate outside a guard band [v_lo, v_hi] of 1e-12 around
// the threshold: below it rejected, above it accepted; inside it, and for
// every candidate whose GC weight is needed, the exact reference arithmetic
// runs.  A NaN v falls in neither range and takes the exact path.
__device__ __forceinline__ double fast_v(const SelectParams& p, int s, double c2, double s2,
                                         double cl2, double sl2) {
  return p.s_cos[s] * c2 * (p.s_cosl[s] * cl2 + p.s_sinl[s] * sl2) + p.s_sin[s] * s2;
}

// Position li of n columns (latitude fastest, L latitudes per longitude) in
// bands of B longitudes visited latitude by latitude: a bijection of [0, n)
// when n is a multiple of L (the last band may be narrower).
__device__ __forceinline__ long long band_index(long long li, long long n, long long B,
                                                long long L) {
  const long long NL = n / L;
  const long long b = li / (B * L), rem = li - b * B * L;
  const long long Bb = min(B, NL - b * B);
  return (b * B + rem % Bb) * L + rem / Bb;
}

template <bool FILL>
__global__ __launch_bounds__(256) void k_select(SelectParams p) {
  __shared__ int s_pre[4][64], s_r0[4][64];
  __shared__ double s_gc[4][64];
  __shared__ int s_q[4][128];            // fill pass: queued candidates (FIFO, site order)
  const int lane = threadIdx.x & 63;
  const int wv = (threadIdx.x >> 6) & 3;
  // Workgroups are dealt round-robin to the 8 XCDs; each XCD walks a
  // contiguous range of visiting positions (the grid is padded to a multiple
  // of 8), and with band_w > 0 the positions go over the columns in bands of
  // longitudes, so the columns in flight on one XCD form a compact patch
  // whose candidate sites share its L2.
  const int ncol = p.h1 - p.h0;
  const int per = (int)gridDim.x / 8;
  const int li = ((int)(blockIdx.x % 8) * per + (int)blockIdx.x / 8) * 4 + wv;
  if (li >= ncol) return;
  const int wid = p.band_w > 0 ? (int)band_index(li, ncol, p.band_w, p.band_l) : li;
  const int ah = p.h0 + wid;
  int sj, ej, si, ei;
  column_center(p.g, ah, &sj, &ej, &si, &ei);
  const double clat = (double)p.g.glat[(sj + ej) / 2 - 1];
  const double clon = (double)p.g.glon[(si + ei) / 2 - 1];
  const double c2 = cos(deg2rad_ref(clat)), s2 = sin(deg2rad_ref(clat));
  const double l2 = deg2rad_ref(clon);
  const double cl2 = cos(l2), sl2 = sin(l2);
  const BinRange R = bin_range(p, clat, clon, p.theta_deg);
  const int b0 = R.b0, b1 = R.b1, nr = R.nr;
  const int* ja = R.ja;
  const int* jb = R.jb;
  long long count = 0;
  long long base = FILL ? p.off[wid] : 0;
  auto exact_gc = [&](int s) {
    const double d = latlon_dist_pre(p.s_cos[s], p.s_sin[s], p.s_lonr[s], c2, s2, l2);
    return gc_weight(d * 1.0 / p.local_dist, p.gc_filter);
  };
  // fill pass: the first n queued candidates (n <= 64, one per lane): exact
  // weights, then their rows written by all 64 lanes -- entry e belongs to the
  // lane whose [pre, pre + rc) holds it (binary search over the wavefront's
  // prefix counts in LDS), so a station column's rows do not serialise on
  // one lane
  auto emit = [&](int n) {
    double gc = 0.0;
    int rc = 0, r0 = 0;
    if (lane < n) {
      const int s = s_q[wv][lane];
      gc = exact_gc(s);
      if (gc > 0.0) {
        rc = p.s_nrow[s];
        r0 = p.s_row0[s];
      }
    }
    const int pre = wave_excl_scan(rc, lane);
    const int tot = __shfl(pre + rc, 63, 64);
    s_pre[wv][lane] = pre;
    s_r0[wv][lane] = r0;
    s_gc[wv][lane] = gc;
    __syncwarp_lds();
    for (int e = lane; e < tot; e += 64) {
      int lo = 0, hi = 63;             // last lane with pre <= e
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[wv][mid] <= e) lo = mid;
        else hi = mid - 1;
      }
      const long long pos = base + e;
      const int row = s_r0[wv][lo] + (e - s_pre[wv][lo]);
      const double g = s_gc[wv][lo];
      if (pos < p.cap) {
        p.row_idx[pos] = p.weights == 2 ? row * p.k : row;
        if (p.weights == 2) {
#pragma clang fp contract(off)
          const double w = (double)p.rinv32[row] * g;
          // w > 0 is a normal number: the hardware square root (1 ulp) suffices
          const f2 sw = {__builtin_amdgcn_sqrtf((float)w), (float)(w * p.dep[row])};
          reinterpret_cast<f2*>(p.row_gc)[pos] = sw;
        } else if (p.weights) {
#pragma clang fp contract(off)
          const double w = (double)p.rinv32[row] * g;
          p.row_gc[pos] = w;
          p.row_wd[pos] = w * p.dep[row];
        } else {
          p.row_gc[pos] = g;
        }
      }
    }
    __syncwarp_lds();
    base += tot;
  };
  int qn = 0;                            // queued candidates (wavefront-uniform)
  for (int b = b0; b <= b1; ++b) {
    for (int r = 0; r < nr; ++r) {
      const int s0 = p.bin_start[b * p.nb_lon + ja[r]];
      const int s1 = p.bin_start[b * p.nb_lon + jb[r] + 1];
      for (int cs = s0; cs < s1; cs += 64) {
        const int s = cs + lane;
        const double v = (s < s1) ? fast_v(p, s, c2, s2, cl2, sl2) : -2.0;
        const bool out = v < p.v_lo;       // sure outside (a NaN is not)
        if (!FILL) {
          const bool in = v > p.v_hi;
          int rc = in ? p.s_nrow[s] : 0;
          const bool band = s < s1 && !out && !in;
          if (__builtin_amdgcn_ballot_w64(band)) {
            if (band && exact_gc(s) > 0.0) rc = p.s_nrow[s];
          }
          count += rc;
        } else {
          const bool keep = s < s1 && !out;
          const unsigned long long m = __builtin_amdgcn_ballot_w64(keep);
          if (keep) {
            const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
            s_q[wv][qn + below] = s;
          }
          qn += __builtin_popcountll(m);
          __syncwarp_lds();
          if (qn >= 64) {
            emit(64);
            // move the remaining qn - 64 (< 64) entries to the queue's head
            const int rest = qn - 64;
            const int t = (lane < rest) ? s_q[wv][64 + lane] : 0;
            __syncwarp_lds();
            if (lane < rest) s_q[wv][lane] = t;
            __syncwarp_lds();
            qn = rest;
          }
        }
      }
    }
  }
  if (FILL && qn > 0) emit(qn);
  if (!FILL) {
    count = wave_sum_ll(count);
    if (lane == 0) p.cnt[wid] = count;
  }
}

// Tile mode: one wavefront per tile of p.tile consecutive analyses of
// [h0, h1).  Lists every row of every site within R_tile of the tile's
// middle centre, R_tile = 2c + the largest distance from the middle centre to
// a centre of the tile (so the union of the analyses' supports is covered, by
// the triangle inequality), in bin order, with the sorted site index.
template <bool FILL>
__global__ __launch_bounds__(256) void k_select_tile(SelectParams p) {
  const int lane = threadIdx.x & 63;
  const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int a0 = p.h0 + t * p.tile;
  if (a0 >= p.h1) return;
  const int a1 = min(a0 + p.tile, p.h1);
  auto centre = [&](int ah, double* la, double* lo) {
    int sj, ej, si, ei;
    column_center(p.g, ah, &sj, &ej, &si, &ei);
    *la = (double)p.g.glat[(sj + ej) / 2 - 1];
    *lo = (double)p.g.glon[(si + ei) / 2 - 1];
  };
  double clat, clon;
  centre((a0 + a1 - 1) / 2, &clat, &clon);
  const double c2 = cos(deg2rad_ref(clat)), s2 = sin(deg2rad_ref(clat));
  const double l2 = deg2rad_ref(clon);
  double spread = 0.0;
  for (int ah = a0; ah < a1; ++ah) {
    double la, lo;
    centre(ah, &la, &lo);
    spread = fmax(spread, latlon_dist_pre(cos(deg2rad_ref(la)), sin(deg2rad_ref(la)),
                                          deg2rad_ref(lo), c2, s2, l2));
  }
  const double r_tile = 2.0 * p.local_dist + spread * (1.0 + 1e-9) + 1e-3;
  const double th = (r_tile / 6371.39306640625) * (180.0 / M_PI) * (1.0 + 1e-6) + 1e-2;
  const BinRange R = bin_range(p, clat, clon, th);
  long long count = 0;
  long long base = FILL ? p.off[t] : 0;
  for (int b = R.b0; b <= R.b1; ++b) {
    for (int r = 0; r < R.nr; ++r) {
      const int s0 = p.bin_start[b * p.nb_lon + R.ja[r]];
      const int s1 = p.bin_start[b * p.nb_lon + R.jb[r] + 1];
      for (int cs = s0; cs < s1; cs += 64) {
        const int s = cs + lane;
        int rc = 0;
        if (s < s1) {
          const double d = latlon_dist_pre(p.s_cos[s], p.s_sin[s], p.s_lonr[s], c2, s2, l2);
          if (d <= r_tile) rc = p.s_nrow[s];
        }
        if (!FILL) {
          count += rc;
        } else {
          const int pre = wave_excl_scan(rc, lane);
          const int tot = __shfl(pre + rc, 63, 64);
          if (rc) {
            const long long pos = base + pre;
            const int r0 = p.s_row0[s];
            for (int u = 0; u < rc; ++u) {
              if (pos + u < p.cap) {
                p.row_idx[pos + u] = r0 + u;
                p.row_site[pos + u] = s;
              }
            }
          }
          base += tot;
        }
      }
    }
  }
  if (!FILL) {
    count = wave_sum_ll(count);
    if (lane == 0) p.cnt[t] = count;
  }
}

// ---------------------------------------------------------------------------
// obs-space products
// ---------------------------------------------------------------------------

// Column profile of the vertical loop: centre coordinate of chunk kc.
__device__ __forceinline__ double chunk_vcoord(const ProdParams& p, int ah, int kc) {
  const int cz = p.cube_nz;
  const int sk = kc * cz + 1;
  const int ek = min(sk + cz - 1, p.nz);
  const int ck = (sk + ek) / 2;
  if (p.vcoord_mode == 0) return (double)ck;
  // ps2alnp (obs_make.F90:418-437) with the hybrid p(ps) of
  // vert_coord_mod.F90:556-563 at the cube centre column
  int sj, ej, si, ei;
  column_center(p.g, ah, &sj, &ej, &si, &ei);
  const int la = (sj + ej) / 2 - 1 + p.halo, lo = (si + ei) / 2 - 1 + p.halo;
  const long long off = ((long long)lo * (p.g.nlat + 2 * p.halo) + la) * p.n;
  const double ps = p.axb[off + p.ps_index];
  const double ph0 = p.hyai[ck - 1] * (p.p0 - p.ptop) + p.hybi[ck - 1] * (ps - p.ptop) + p.ptop;
  const double ph1 = p.hyai[ck] * (p.p0 - p.ptop) + p.hybi[ck] * (ps - p.ptop) + p.ptop;
  return log(0.5 * (ph0 + ph1));
}

// weight of one local obs row: rdiaginv (real(4)) * rloc, times the
// vertical GC factor (vertical_localization, letkf_yb.F90:240-260)
__device__ __forceinline__ double row_weight(const ProdParams& p, int row, double gc,
                                             double vc) {
#pragma clang fp contract(off)
  double w = (double)p.rinv32[row] * gc;
  if (p.vertical) {
    double ri = fabs(vc - p.vco[row]) / p.vert_dist;
    w = w * gc_weight(ri, p.gc_filter_vertical);
  }
  return w;
}

// Workgroups are dealt round-robin to the 8 XCDs; remap them so that each
// XCD walks a contiguous range of logical blocks (neighbouring analyses share
// obs rows in that XCD's L2).  The grid is padded to a multiple of 8.
__device__ __forceinline__ int xcd_block(int n_blocks) {
  const int per = (n_blocks + 7) / 8;
  return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

// Batch-local analysis visited at position li.  With band_w > 0 the
// analyses (one per column, columns latitude-fastest) are visited in bands of
// band_w longitudes, latitude by latitude, so the analyses in flight on an
// XCD form a compact patch and share more of their gathered obs rows in its
// L2 than a run along one meridian does.  A bijection of [0, n_an): the last
// band may be narrower.
__device__ __forceinline__ long long analysis_at(const ProdParams& p, long long li) {
  if (p.order) return p.order[li];
  if (p.band_w <= 0) return li;
  return band_index(li, p.n_an, p.band_w, p.band_l);
}

// The analyses a products launch visits.  WPA is 1 (a wavefront per
// analysis) or 4 (the whole 256-lane workgroup); body(li) runs the products
// of batch-local analysis li.  SCHED (separate instantiations, so that each
// schedule keeps its own register budget and timing costs nothing when off):
//   0  plain: the analysis at the group's logical position (column order,
//      or bands of longitudes);
//   1  dispatch order over a bucket segment: the launch's groups map to the
//      positions [seg_off[seg_lo], seg_off[seg_hi]) of the order list
//      (length buckets, longest first, grid order within a bucket), dealt to
//      the 8 XCDs in contiguous eighths, and the hardware dispatcher hands
//      them to CUs as CUs free up -- the longest analyses start first;
//   3  persistent work queue over a bucket segment: as many workgroups as
//      fit at once, the groups take positions from a counter per XCD eighth
//      (then help the other eighths) until the segment is used up;
//   2, 4  as 1, 3 with timing: each analysis adds its wall-clock ticks to its
//      bucket position, and the launch's four totals get the earliest start,
//      the latest end, the summed busy ticks and the analyses (1, 2) or
//      workers (3, 4) that ran.
// Within a bucket segment the analyses an XCD runs are neighbours in the
// grid and share their obs rows in that XCD's L2.
template <int WPA, int SCHED, class Body>
__device__ __forceinline__ void for_analyses(const ProdParams& p, int lane, Body body) {
  static_assert(WPA == 1 || WPA == 4, "a wavefront or a workgroup per analysis");
  const int wave = threadIdx.x >> 6;
  const int apb = (blockDim.x >> 6) / WPA;
  if constexpr (SCHED == 0) {
    const long long li = (long long)xcd_block(p.n_blocks) * apb + wave / WPA;
    if (li >= p.n_an) return;
    body(analysis_at(p, li));
    return;
  }
  constexpr bool timing = SCHED == 2 || SCHED == 4;
  const bool leader = lane == 0 && (WPA == 1 || wave == 0);
  const long long lo = p.seg_off[p.seg_lo], hi = p.seg_off[p.seg_hi], len = hi - lo;
  auto bucket_of_pos = [&](long long pos) {
    int j = p.seg_lo;
    while (j + 1 < p.seg_hi && pos >= p.seg_off[j + 1]) ++j;
    return j;
  };
  if constexpr (SCHED == 1 || SCHED == 2) {
    // Workgroups are dealt round-robin to the 8 XCDs by index and dispatched
    // in index order.  XCD x = blockIdx % 8 runs the x-th eighth of every
    // bucket of the segment, bucket after bucket: all XCDs work on the same
    // bucket at the same time (longest first, balanced), and each on a
    // contiguous piece of it, grid neighbours whose obs rows its L2 holds.
    const int x = blockIdx.x & 7;
    long long q = (long long)(blockIdx.x >> 3) * apb + wave / WPA;   // index in x's share
    long long pos = -1;
    for (int j = p.seg_lo; j < p.seg_hi; ++j) {
      const long long b0 = p.seg_off[j], nb = p.seg_off[j + 1] - b0;
      const long long e0 = b0 + nb * x / 8, e1 = b0 + nb * (x + 1) / 8;
      if (q < e1 - e0) {
        pos = e0 + q;
        break;
      }
      q -= e1 - e0;
    }
    if (pos < 0) return;
    unsigned long long t0 = 0;
    if constexpr (timing) t0 = wall_clock64();
    body((long long)__builtin_amdgcn_readfirstlane(p.order[pos]));
    if constexpr (timing) {
      const unsigned long long t1 = wall_clock64();
      if (leader) {
        atomicAdd(&p.bucket_ticks[bucket_of_pos(pos)], t1 - t0);
        atomicMin(&p.launch_ticks[0], t0);
        atomicMax(&p.launch_ticks[1], t1);
        atomicAdd(&p.launch_ticks[2], t1 - t0);
        atomicAdd(&p.launch_ticks[3], 1ull);
      }
    }
    return;
  }
  unsigned long long busy = 0, t_first = ~0ull, t_last = 0;
  __shared__ int s_next;
  // kQueueSplit: one counter per XCD eighth of the segment (each XCD works
  // on its own region); otherwise one counter, all XCDs at the queue's front
  int vx = kQueueSplit ? (blockIdx.x & 7) : 0, tried = kQueueSplit ? 0 : 7;
  for (;;) {
    int v;
    if constexpr (WPA == 1) {
      v = 0;
      if (lane == 0) v = atomicAdd(p.q_counter + vx, 1);
      v = __builtin_amdgcn_readfirstlane(__shfl(v, 0, 64));
    } else {
      if (threadIdx.x == 0) s_next = atomicAdd(p.q_counter + vx, 1);
      __syncthreads();
      v = __builtin_amdgcn_readfirstlane(s_next);
      __syncthreads();
    }
    const long long c0 = kQueueSplit ? lo + len * vx / 8 : lo;
    const long long c1 = kQueueSplit ? lo + len * (vx + 1) / 8 : hi;
    const long long pos = c0 + v;
    if (pos >= c1) {
      if (++tried == 8) break;
      vx = (vx + 1) & 7;
      continue;
    }
    unsigned long long t0 = 0;
    if constexpr (timing) t0 = wall_clock64();
    // the analysis is uniform over the wavefront: keep it (and every address
    // derived from it) in scalar registers
    body((long long)__builtin_amdgcn_readfirstlane(p.order[pos]));
    if constexpr (timing) {
      const unsigned long long t1 = wall_clock64();
      busy += t1 - t0;
      t_first = t0 < t_first ? t0 : t_first;
      t_last = t1;
      if (leader) atomicAdd(&p.bucket_ticks[bucket_of_pos(pos)], t1 - t0);
    }
  }
  if constexpr (timing) {
    if (leader) {
      if (busy > 0) {
        atomicMin(&p.launch_ticks[0], t_first);
        atomicMax(&p.launch_ticks[1], t_last);
        atomicAdd(&p.launch_ticks[2], busy);
      }
      atomicAdd(&p.launch_ticks[3], 1ull);
    }
  }
}

// Reduced-precision products: row lists longer than kFlushMinRows add their
// FP32 partial sums to FP64 every kFlushRows rows.
constexpr int kFlushMinRows = 3072;
constexpr int kFlushRows = 1024;

template <int T>
struct PairTable {
  static constexpr int N = T * (T + 1) / 2;
  __device__ static constexpr int I(int pid) {
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= pid) ++i;
    return i;
  }
  __device__ static constexpr int J(int pid) { return pid - I(pid) * (I(pid) + 1) / 2; }
};

template <int KP, int WS, int PPW, bool PW>
__device__ __forceinline__ void products_f64_wave(const ProdParams& p, long long a, int ah,
                                                  double vc, long long r0, long long r1,
                                                  int lane) {
  constexpr int T = KP / 16;
  using PT = PairTable<T>;
  constexpr int P0 = WS * PPW;
  constexpr int P1 = (P0 + PPW < PT::N) ? P0 + PPW : PT::N;
  constexpr int NP = P1 - P0;
  const int i16 = lane & 15, q = lane >> 4;
  d4 acc[NP];
#pragma unroll
  for (int x = 0; x < NP; ++x) acc[x] = d4{0.0, 0.0, 0.0, 0.0};
  double bacc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) bacc[t] = 0.0;
  // G groups of 4 rows per step (lane q takes rows 4g + q), loaded one step
  // ahead of the MFMAs that use them; positions past the end of the row list
  // read its last row with zero weight (every load in bounds, no branch).
  // More groups cost wavefronts per SIMD, which the per-row vertical weights
  // of the k=32 level loop need more
  constexpr int G = KP == 64 ? 2 : 1;
  const int n = (int)(r1 - r0);
  auto load = [&](int rl, double* y, double& w, double& wd) {
    const int rc = rl < n ? rl : n - 1;
    const int row = p.row_idx[r0 + rc];
    double wv, wdv;
    if constexpr (PW) {
      wv = p.row_gc[r0 + rc];
      wdv = p.row_wd[r0 + rc];
    } else {
      wv = row_weight(p, row, p.row_gc[r0 + rc], vc);
      wdv = wv * p.dep[row];
    }
    const bool in = rl < n;
    w = in ? wv : 0.0;
    wd = in ? wdv : 0.0;
    const double* yr = p.yb64 + (long long)row * p.k;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int m = 16 * t + i16;
      // an unconditional load of an in-row member and a select: no branch
      const double yl = yr[m < p.k ? m : p.k - 1];
      y[t] = (m < p.k) ? yl : 0.0;
    }
  };
  double yn[G][T], wn[G], wdn[G];
  if (n > 0) {
#pragma unroll
    for (int g = 0; g < G; ++g) load(4 * g + q, yn[g], wn[g], wdn[g]);
  }
  for (int g0 = 0; g0 < n; g0 += 4 * G) {
    double y[G][T], w[G], wd[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int t = 0; t < T; ++t) y[g][t] = yn[g][t];
      w[g] = wn[g];
      wd[g] = wdn[g];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) load(g0 + 4 * G + 4 * g + q, yn[g], wn[g], wdn[g]);
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int x = 0; x < NP; ++x) {
        const int I = PT::I(P0 + x), J = PT::J(P0 + x);
        acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[g][I], w[g] * y[g][J], acc[x], 0, 0, 0);
      }
      if (WS == 0) {
#pragma unroll
        for (int t = 0; t < T; ++t) bacc[t] = fma(wd[g], y[g][t], bacc[t]);
      }
    }
  }
  double* A = p.A + (a - p.a0) * (long long)KP * KP;
#pragma unroll
  for (int x = 0; x < NP; ++x) {
    const int I = PT::I(P0 + x), J = PT::J(P0 + x);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * I + q + 4 * r, col = 16 * J + i16;
      A[(long long)col * KP + row] = acc[x][r];
      if (I != J) A[(long long)row * KP + col] = acc[x][r];
    }
  }
  if (WS == 0) {
    double* b = p.b + (a - p.a0) * KP;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      double v = bacc[t];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (q == 0) b[16 * t + i16] = v;
    }
  }
}

template <int KP, int SCHED, bool PW>
__global__ __launch_bounds__(256) void k_products_f64(ProdParams p) {
  constexpr int T = KP / 16;
  constexpr int NPAIR = T * (T + 1) / 2;
  constexpr int WPA = (NPAIR + 9) / 10;
  constexpr int PPW = (NPAIR + WPA - 1) / WPA;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wsub = wave % WPA;
  for_analyses<WPA, SCHED>(p, lane, [&](long long li) {
    const long long a = p.a0 + li;
    const int ah = (int)(a / p.nzc), kc = (int)(a % p.nzc);
    // row lists past the list buffer's capacity (the call then reruns with a
    // larger one) are cut at it, so no read leaves the buffer
    const long long r0 = min(p.row_off[ah - p.h0], p.row_cap);
    const long long r1 = min(p.row_off[ah - p.h0 + 1], p.row_cap);
    const double vc = p.vertical ? chunk_vcoord(p, ah, kc) : 0.0;
    if (wsub == 0) products_f64_wave<KP, 0, PPW, PW>(p, a, ah, vc, r0, r1, lane);
    if constexpr (WPA > 1) {
      if (wsub == 1) products_f64_wave<KP, 1, PPW, PW>(p, a, ah, vc, r0, r1, lane);
      if (wsub == 2) products_f64_wave<KP, 2, PPW, PW>(p, a, ah, vc, r0, r1, lane);
      if (wsub == 3) products_f64_wave<KP, 3, PPW, PW>(p, a, ah, vc, r0, r1, lane);
    }
  });
}

template <int KP, int WS, int PPW, bool PW>
__device__ __forceinline__ void products_f32_wave(const ProdParams& p, long long a, int ah,
                                                  double vc, long long r0, long long r1,
                                                  int lane) {
  constexpr int T = KP / 16;
  using PT = PairTable<T>;
  constexpr int P0 = WS * PPW;
  constexpr int P1 = (P0 + PPW < PT::N) ? P0 + PPW : PT::N;
  constexpr int NP = P1 - P0;
  const int i16 = lane & 15, q = lane >> 4;
  f4 acc[NP];
#pragma unroll
  for (int x = 0; x < NP; ++x) acc[x] = f4{0.f, 0.f, 0.f, 0.f};
  float bacc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) bacc[t] = 0.f;
  // row rl of the analysis' list; past the end the list's last row is read
  // (every load in bounds without a branch) with zero weight
  const int n = (int)(r1 - r0);
  auto load = [&](int rl, float* y, float& wf, float& wdf) {
    const int rc = rl < n ? rl : n - 1;
    const int row = p.row_idx[r0 + rc];
    double w, wd;
    if constexpr (PW) {
      w = p.row_gc[r0 + rc];
      wd = p.row_wd[r0 + rc];
    } else {
      w = row_weight(p, row, p.row_gc[r0 + rc], vc);
      wd = w * p.dep[row];
    }
    const bool in = rl < n;
    wf = in ? (float)w : 0.f;
    wdf = in ? (float)wd : 0.f;
    const float* yr = p.yb32 + (long long)row * p.k;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int m = 16 * t + i16;
      const float yl = yr[m < p.k ? m : p.k - 1];
      y[t] = (m < p.k) ? yl : 0.f;
    }
  };
  // G groups of 4 rows per step (lane q takes rows 4g + q), loaded one step
  // ahead of the MFMAs that use them
  constexpr int G = KP == 64 ? 4 : 1;
  float yn[G][T], wfn[G], wdn[G];
  if (n > 0) {
#pragma unroll
    for (int g = 0; g < G; ++g) load(4 * g + q, yn[g], wfn[g], wdn[g]);
  }
  // Long row lists: the FP32 sums of every kFlushRows rows are added to FP64
  // (A in memory, b in registers), so the FP32 accumulation error grows with
  // kFlushRows instead of p
  double* A = p.A + (a - p.a0) * (long long)KP * KP;
  const bool chunked = n > kFlushMinRows;
  bool flushed = false;
  double bsum[T];
#pragma unroll
  for (int t = 0; t < T; ++t) bsum[t] = 0.0;
  for (int g0 = 0; g0 < n; g0 += 4 * G) {
    if (chunked && g0 > 0 && g0 % kFlushRows == 0) {
#pragma unroll
      for (int x = 0; x < NP; ++x) {
        const int I = PT::I(P0 + x), J = PT::J(P0 + x);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double* e = A + (long long)(16 * J + i16) * KP + 16 * I + 4 * q + r;
          *e = (flushed ? *e : 0.0) + (double)acc[x][r];
        }
        acc[x] = f4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int t = 0; t < T; ++t) {
        bsum[t] += (double)bacc[t];
        bacc[t] = 0.f;
      }
      flushed = true;
    }
    float y[G][T], wf[G], wd[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int t = 0; t < T; ++t) y[g][t] = yn[g][t];
      wf[g] = wfn[g];
      wd[g] = wdn[g];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) load(g0 + 4 * G + 4 * g + q, yn[g], wfn[g], wdn[g]);
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int x = 0; x < NP; ++x) {
        const int I = PT::I(P0 + x), J = PT::J(P0 + x);
        acc[x] = __builtin_amdgcn_mfma_f32_16x16x4f32(y[g][I], wf[g] * y[g][J], acc[x], 0, 0, 0);
      }
      if (WS == 0) {
#pragma unroll
        for (int t = 0; t < T; ++t) bacc[t] = fmaf(wd[g], y[g][t], bacc[t]);
      }
    }
  }
#pragma unroll
  for (int x = 0; x < NP; ++x) {
    const int I = PT::I(P0 + x), J = PT::J(P0 + x);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * I + 4 * q + r, col = 16 * J + i16;
      const double v = (flushed ? A[(long long)col * KP + row] : 0.0) + (double)acc[x][r];
      A[(long long)col * KP + row] = v;
      if (I != J) A[(long long)row * KP + col] = v;
    }
  }
  if (WS == 0) {
    double* b = p.b + (a - p.a0) * KP;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      double v = bsum[t] + (double)bacc[t];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (q == 0) b[16 * t + i16] = v;
    }
  }
}

template <int KP, int SCHED, bool PW>
__global__ __launch_bounds__(256) void k_products_f32(ProdParams p) {
  constexpr int T = KP / 16;
  constexpr int NPAIR = T * (T + 1) / 2;
  constexpr int WPA = (NPAIR + 9) / 10;
  constexpr int PPW = (NPAIR + WPA - 1) / WPA;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wsub = wave % WPA;
  for_analyses<WPA, SCHED>(p, lane, [&](long long li) {
    const long long a = p.a0 + li;
    const int ah = (int)(a / p.nzc), kc = (int)(a % p.nzc);
    // row lists past the list buffer's capacity (the call then reruns with a
    // larger one) are cut at it, so no read leaves the buffer
    const long long r0 = min(p.row_off[ah - p.h0], p.row_cap);
    const long long r1 = min(p.row_off[ah - p.h0 + 1], p.row_cap);
    const double vc = p.vertical ? chunk_vcoord(p, ah, kc) : 0.0;
    if (wsub == 0) products_f32_wave<KP, 0, PPW, PW>(p, a, ah, vc, r0, r1, lane);
    if constexpr (WPA > 1) {
      if (wsub == 1) products_f32_wave<KP, 1, PPW, PW>(p, a, ah, vc, r0, r1, lane);
      if (wsub == 2) products_f32_wave<KP, 2, PPW, PW>(p, a, ah, vc, r0, r1, lane);
      if (wsub == 3) products_f32_wave<KP, 3, PPW, PW>(p, a, ah, vc, r0, r1, lane);
    }
  });
}

// BF16 split of yhat = sqrt(w) y (FP32), FP32 accumulation (b = sum w d y
// too).  k <= 64: two terms, hi = bf16(yhat), lo = bf16(yhat - hi) (16 bits),
// A ~ hi hi^T + hi lo^T + lo hi^T.  k = 65..128, where the increment's
// sensitivity to A grows with the eigenvalue spread: three terms
// (hi, mid, lo: 24 bits) and six products (products_bf16_shared).

// One 32-row group position of an analysis' row list (rl, with n rows):
// sqrt(w) and w d in FP32, zero past the end of the list (whose last row is
// read in its place, so every load stays in bounds without a branch).
// TAIL false: the caller knows rl < n.
template <bool PW, bool TAIL = true>
__device__ __forceinline__ void group_row(const ProdParams& p, long long r0, int n, int rl,
                                          int row, double vc, float& s, float& wdf) {
  const bool in = !TAIL || rl < n;
  if constexpr (PW) {
    // the fill pass stored {sqrt(w), w d} in FP32; the lists are padded
    // (kListPad): read in place, the select drops positions past the end
    const f2 sw = reinterpret_cast<const f2*>(p.row_gc)[r0 + rl];
    s = in ? sw.x : 0.f;
    wdf = in ? sw.y : 0.f;
  } else {
    const int rc = in ? rl : n - 1;
    const double w = row_weight(p, row, p.row_gc[r0 + rc], vc);
    const double wd = w * p.dep[row];
    // w > 0 is a normal number: the hardware square root (1 ulp) suffices
    s = in ? __builtin_amdgcn_sqrtf((float)w) : 0.f;
    wdf = in ? (float)wd : 0.f;
  }
}

template <int KP, int WS, int PPW, int G, bool PW>
__device__ __forceinline__ void products_bf16_wave(const ProdParams& p, long long a, int ah,
                                                   double vc, long long r0, long long r1,
                                                   int lane) {
  constexpr int T = KP / 16;
  using PT = PairTable<T>;
  constexpr int P0 = WS * PPW;
  constexpr int P1 = (P0 + PPW < PT::N) ? P0 + PPW : PT::N;
  constexpr int NP = P1 - P0;
  const int i16 = lane & 15, q = lane >> 4;
  f4 acc[NP];
#pragma unroll
  for (int x = 0; x < NP; ++x) acc[x] = f4{0.f, 0.f, 0.f, 0.f};
  float bacc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) bacc[t] = 0.f;
  const int n = (int)(r1 - r0);
  const int* __restrict__ ridx = p.row_idx + r0;
  // G groups of 32 rows per step (more gathers in flight for G = 2; the
  // MFMAs and the b sums run in the same order, so the sums are the same
  // bits for every G).  The row indices of the next step are loaded one step
  // ahead, so only the Yb gathers wait on memory.  The lists are padded past
  // their end with valid row indices (kListPad), so no index is clamped.
  // Yb is read through a buffer resource at a 32-bit byte offset per row
  // (the list holds row * k when pre-weighted), one 4T-byte load per lane and
  // row: lane i16 holds members T i16 .. T i16 + T - 1, so tile t of the
  // MFMAs is the members m = T i + t (i = 0..15), and A and b are stored
  // with the members put back in order.  Members m >= k read the next row
  // (or 0 past the resource) and are dropped when A and b are stored.
  // Positions past the end of the list (last step only) load from an offset
  // past the resource, which reads 0, and enter with weight 0: they add
  // exact zeros whatever the rows that follow in the list hold.
  int idx_n[G][8];
#pragma unroll
  for (int gg = 0; gg < G; ++gg)
#pragma unroll
    for (int u = 0; u < 8; ++u) idx_n[gg][u] = ridx[32 * gg + 8 * q + u];
  const __amdgpu_buffer_rsrc_t yrs =
      __builtin_amdgcn_make_buffer_rsrc((void*)p.yb32, 0, (int)p.yb32_bytes, 0x00020000);
  static_assert(T == 2 || T == 4, "one 8- or 16-byte load per lane and row");
  const int voff = 4 * T * i16;
  const int rowb = PW ? 4 : 4 * p.k;           // bytes per list entry unit
  // one step of G groups from position g0; TAIL: some positions may lie past
  // the end of the list (the last step only)
  auto step = [&](int g0, auto tail) {
    constexpr bool TAIL = decltype(tail)::value;
    bf8 hi[G][T], lo[G][T];
    int idx[G][8];
#pragma unroll
    for (int gg = 0; gg < G; ++gg)
#pragma unroll
      for (int u = 0; u < 8; ++u) idx[gg][u] = idx_n[gg][u];
    if constexpr (!TAIL) {
#pragma unroll
      for (int gg = 0; gg < G; ++gg)
#pragma unroll
        for (int u = 0; u < 8; ++u)
          idx_n[gg][u] = ridx[g0 + 32 * G + 32 * gg + 8 * q + u];
    }
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
      float s[8], wdf[8], yv[8][T];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        group_row<PW, TAIL>(p, r0, n, g0 + 32 * gg + 8 * q + u, idx[gg][u], vc, s[u], wdf[u]);
        const int rl = g0 + 32 * gg + 8 * q + u;
        const int yo = (!TAIL || rl < n) ? idx[gg][u] * rowb + voff : (int)kYbOffsetLimit;
        if constexpr (T == 4) {
          const f4 v = raw_buffer_load_f4(yrs, yo, 0, 0);
#pragma unroll
          for (int t = 0; t < T; ++t) yv[u][t] = v[t];
        } else {
          const f2 v = raw_buffer_load_f2(yrs, yo, 0, 0);
#pragma unroll
          for (int t = 0; t < T; ++t) yv[u][t] = v[t];
        }
      }
      // rows u and u + 1 convert as one packed pair: hi = bf16(s y) and
      // lo = bf16(s y - hi), both rounded to nearest even
#pragma unroll
      for (int t = 0; t < T; ++t) {
        u4 hw, lw;
#pragma unroll
        for (int up = 0; up < 4; ++up) {
          const f2 yh = {s[2 * up] * yv[2 * up][t], s[2 * up + 1] * yv[2 * up + 1][t]};
          const bf2 h2 = __builtin_convertvector(yh, bf2);
          const bf2 l2 = __builtin_convertvector(yh - __builtin_convertvector(h2, f2), bf2);
          hw[up] = __builtin_bit_cast(unsigned, h2);
          lw[up] = __builtin_bit_cast(unsigned, l2);
        }
        hi[gg][t] = __builtin_bit_cast(bf8, hw);
        lo[gg][t] = __builtin_bit_cast(bf8, lw);
        if (WS == 0) {
#pragma unroll
          for (int u = 0; u < 8; ++u) bacc[t] = fmaf(wdf[u], yv[u][t], bacc[t]);
        }
      }
    }
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
#pragma unroll
      for (int x = 0; x < NP; ++x) {
        const int I = PT::I(P0 + x), J = PT::J(P0 + x);
        acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[gg][I], hi[gg][J], acc[x], 0, 0, 0);
        acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[gg][I], lo[gg][J], acc[x], 0, 0, 0);
        acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[gg][I], hi[gg][J], acc[x], 0, 0, 0);
      }
    }
  };
  const int n_full = n / (32 * G) * (32 * G);
  for (int g0 = 0; g0 < n_full; g0 += 32 * G) step(g0, std::false_type{});
  if (n_full < n) step(n_full, std::true_type{});
  // A in FP32 tile order (a_entry): tile entry (4q + r, i16) of tiles (I, J)
  // is members T (4q + r) + I and T i16 + J, stored at positions
  // 16 I + 4q + r and 16 J + i16, so the sixteen lanes of a quarter write
  // consecutive words
  float* A = reinterpret_cast<float*>(p.A + (a - p.a0) * (long long)KP * KP);
#pragma unroll
  for (int x = 0; x < NP; ++x) {
    const int I = PT::I(P0 + x), J = PT::J(P0 + x);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mr = T * (4 * q + r) + I, mc = T * i16 + J;
      const int pr = 16 * I + 4 * q + r, pc = 16 * J + i16;
      const float v = (mr < p.k && mc < p.k) ? acc[x][r] : 0.f;
      A[pc * KP + pr] = v;
      if (I != J) A[pr * KP + pc] = v;
    }
  }
  if (WS == 0) {
    double* b = p.b + (a - p.a0) * KP;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      float v = bacc[t];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (q == 0) b[T * i16 + t] = (T * i16 + t < p.k) ? (double)v : 0.0;
    }
  }
}

// k = 65..128 (four wavefronts per analysis): each 32-row group is converted
// once.  Wavefront WS splits the rows of its T/4 column tiles into BF16 hi/lo
// (and accumulates their part of b), the block shares the operands through a
// double-buffered LDS stage with one barrier per group, and each wavefront
// then runs the MFMAs of its tile pairs from LDS.  Same values and
// accumulation order as products_bf16_wave.
template <int KP, int WS, int PPW, bool PW>
__device__ __forceinline__ void products_bf16_shared(const ProdParams& p, long long a, int ah,
                                                     double vc, long long r0, long long r1,
                                                     int lane, bf8* __restrict__ sh) {
  constexpr int T = KP / 16;
  constexpr int TW = T / 4;                  // tiles converted by this wavefront
  using PT = PairTable<T>;
  constexpr int P0 = WS * PPW;
  constexpr int P1 = (P0 + PPW < PT::N) ? P0 + PPW : PT::N;
  constexpr int NP = P1 - P0;
  const int i16 = lane & 15, q = lane >> 4;
  f4 acc[NP];
#pragma unroll
  for (int x = 0; x < NP; ++x) acc[x] = f4{0.f, 0.f, 0.f, 0.f};
  float bacc[TW];
#pragma unroll
  for (int t = 0; t < TW; ++t) bacc[t] = 0.f;
  const int n = (int)(r1 - r0);
  const int* __restrict__ ridx = p.row_idx + r0;
  int idx_n[8];
  if (n > 0) {
#pragma unroll
    for (int u = 0; u < 8; ++u) idx_n[u] = ridx[min(8 * q + u, n - 1)];
  }
  // long lists add their FP32 partial sums to FP64 every kFlushRows rows
  // (A in memory, b in registers)
  double* A = p.A + (a - p.a0) * (long long)KP * KP;
  const bool chunked = n > kFlushMinRows;
  bool flushed = false;
  double bsum[TW];
#pragma unroll
  for (int t = 0; t < TW; ++t) bsum[t] = 0.0;
  int buf = 0;
  for (int g0 = 0; g0 < n; g0 += 32, buf ^= 1) {
    if (chunked && g0 > 0 && g0 % kFlushRows == 0) {
#pragma unroll
      for (int x = 0; x < NP; ++x) {
        const int I = PT::I(P0 + x), J = PT::J(P0 + x);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double* e = A + (long long)(16 * J + i16) * KP + 16 * I + 4 * q + r;
          *e = (flushed ? *e : 0.0) + (double)acc[x][r];
        }
        acc[x] = f4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        bsum[t] += (double)bacc[t];
        bacc[t] = 0.f;
      }
      flushed = true;
    }
    bf8 hi[TW], mi[TW], lo[TW];
    int idx[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) idx[u] = idx_n[u];
#pragma unroll
    for (int u = 0; u < 8; ++u) idx_n[u] = ridx[min(g0 + 32 + 8 * q + u, n - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      float s, wdf;
      group_row<PW>(p, r0, n, g0 + 8 * q + u, idx[u], vc, s, wdf);
      // pre-weighted lists hold the row's offset row * k
      const float* yr = p.yb32 + (PW ? (long long)idx[u] : (long long)idx[u] * p.k);
#pragma unroll
      for (int tt = 0; tt < TW; ++tt) {
        const int m = 16 * (WS * TW + tt) + i16;
        const float yl = yr[m < p.k ? m : p.k - 1];     // a select, not a branch
        const float yv = (m < p.k) ? yl : 0.f;
        // yh = hi + mid + lo: 24 bits in three BF16 terms (each difference
        // is exact in FP32)
        const float yh = s * yv;
        const __bf16 h = (__bf16)yh;
        const float r1v = yh - (float)h;
        const __bf16 m1 = (__bf16)r1v;
        hi[tt][u] = h;
        mi[tt][u] = m1;
        lo[tt][u] = (__bf16)(r1v - (float)m1);
        bacc[tt] = fmaf(wdf, yv, bacc[tt]);
      }
    }
    bf8* sb = sh + buf * (3 * T * 64);
#pragma unroll
    for (int tt = 0; tt < TW; ++tt) {
      sb[(WS * TW + tt) * 64 + lane] = hi[tt];
      sb[(T + WS * TW + tt) * 64 + lane] = mi[tt];
      sb[(2 * T + WS * TW + tt) * 64 + lane] = lo[tt];
    }
    __syncthreads();
    // the buffer written next is the other one; it was last read before this
    // barrier, so one barrier per group suffices.  The six products that
    // carry 24 bits -- hi.hi, hi.mid, mid.hi, mid.mid, hi.lo, lo.hi -- run in
    // two rounds, the second reading lo into mid's registers
    bf8 H[T], M[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      H[t] = sb[t * 64 + lane];
      M[t] = sb[(T + t) * 64 + lane];
    }
#pragma unroll
    for (int x = 0; x < NP; ++x) {
      const int I = PT::I(P0 + x), J = PT::J(P0 + x);
      acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H[I], H[J], acc[x], 0, 0, 0);
      acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H[I], M[J], acc[x], 0, 0, 0);
      acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(M[I], H[J], acc[x], 0, 0, 0);
      acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(M[I], M[J], acc[x], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) M[t] = sb[(2 * T + t) * 64 + lane];
#pragma unroll
    for (int x = 0; x < NP; ++x) {
      const int I = PT::I(P0 + x), J = PT::J(P0 + x);
      acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H[I], M[J], acc[x], 0, 0, 0);
      acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(M[I], H[J], acc[x], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < NP; ++x) {
    const int I = PT::I(P0 + x), J = PT::J(P0 + x);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * I + 4 * q + r, col = 16 * J + i16;
      const double v = (flushed ? A[(long long)col * KP + row] : 0.0) + (double)acc[x][r];
      A[(long long)col * KP + row] = v;
      if (I != J) A[(long long)row * KP + col] = v;
    }
  }
  double* b = p.b + (a - p.a0) * KP;
#pragma unroll
  for (int tt = 0; tt < TW; ++tt) {
    double v = bsum[tt] + (double)bacc[tt];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (q == 0) b[16 * (WS * TW + tt) + i16] = v;
  }
}

// Three wavefronts per SIMD, and no more: given room for a fourth, the
// compiler serialises the row gathers to fit its registers.  The persistent
// queue loops (SCHED 3, 4) keep their state across analyses and get two.
template <int KP, int G, int SCHED, bool PW>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(SCHED >= 3 ? 2 : 3, SCHED >= 3 ? 2 : 3)))
void k_products_bf16(ProdParams p) {
  constexpr int T = KP / 16;
  constexpr int NPAIR = T * (T + 1) / 2;
  constexpr int WPA = (NPAIR + 9) / 10;
  constexpr int PPW = (NPAIR + WPA - 1) / WPA;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wsub = wave % WPA;
  if constexpr (WPA == 4) static_assert(KP % 64 == 0, "four wavefronts share an analysis");
  __shared__ bf8 sh[WPA == 1 ? 1 : 2 * 3 * (KP / 16) * 64];
  for_analyses<WPA, SCHED>(p, lane, [&](long long li) {
    const long long a = p.a0 + li;
    const int ah = (int)(a / p.nzc), kc = (int)(a % p.nzc);
    // row lists past the list buffer's capacity (the call then reruns with a
    // larger one) are cut at it, so no read leaves the buffer
    const long long r0 = min(p.row_off[ah - p.h0], p.row_cap);
    const long long r1 = min(p.row_off[ah - p.h0 + 1], p.row_cap);
    const double vc = p.vertical ? chunk_vcoord(p, ah, kc) : 0.0;
    if constexpr (WPA == 1) {
      products_bf16_wave<KP, 0, PPW, G, PW>(p, a, ah, vc, r0, r1, lane);
    } else {
      if (wsub == 0) products_bf16_shared<KP, 0, PPW, PW>(p, a, ah, vc, r0, r1, lane, sh);
      if (wsub == 1) products_bf16_shared<KP, 1, PPW, PW>(p, a, ah, vc, r0, r1, lane, sh);
      if (wsub == 2) products_bf16_shared<KP, 2, PPW, PW>(p, a, ah, vc, r0, r1, lane, sh);
      if (wsub == 3) products_bf16_shared<KP, 3, PPW, PW>(p, a, ah, vc, r0, r1, lane, sh);
      // the next analysis' first stage must not overwrite operands another
      // wavefront still reads
      __syncthreads();
    }
  });
}

// ---------------------------------------------------------------------------
// transform: FP64 eigensolve + Pa, T, overline_w, trans, fused update
// ---------------------------------------------------------------------------
//
// One wavefront per analysis, lane r owns row r.  The symmetric matrix and the
// eigenvector matrix live in LDS (column-major, leading dimension KMAX+1):
//   1. Householder tridiagonalisation (Golub & Van Loan Alg. 8.3.1)
//   2. explicit Q from the stored reflectors (backward accumulation)
//   3. implicit QL with Wilkinson shift on the tridiagonal (Bowdler, Martin,
//      Reinsch & Wilkinson 1968), rotations applied to Z = Q in place
//   4. lambda_j += (k-1)/rho (letkf_core.F90:110-113), guard for reduced modes
//   5. overline_w = Z diag(1/lambda) Z^T b, T = Z diag(sqrt((k-1)/lambda)) Z^T
//   6. trans = T + overline_w 1^T; xa(:,v) = trans^T xb(:,v) + axb(v), then
//      axa(v) = mean_e xa(e,v), xa(:,v) -= axa(v)

template <int KMAX>
__global__ __launch_bounds__(64) void k_transform(XformParams p) {
  constexpr int LD = KMAX + 1;
  __shared__ double sA[KMAX * LD];
  __shared__ double sZ[KMAX * LD];
  __shared__ double sv[KMAX], sw[KMAX], sd[KMAX + 1], se[KMAX + 1], stau[KMAX], sb[KMAX],
      st[KMAX], sx[KMAX];
  const int lane = threadIdx.x;
  const long long a = p.list ? (long long)p.list[blockIdx.x] : p.a0 + blockIdx.x;
  if (!p.list && a >= p.a0 + p.n_an) return;
  const int k = p.k;
  const bool act_r = lane < k;
  const double* Ag = p.A + (a - p.a0) * (long long)p.kp * p.kp;
  const double* bg = p.b + (a - p.a0) * p.kp;

  // ---- load lower triangle, symmetrise, zero padding ----
  for (int c = 0; c < KMAX; ++c) {
    double v = 0.0;
    if (lane < KMAX) {
      if (act_r && c < k)
        v = (lane >= c) ? a_entry(Ag, p.a_f32, p.kp, lane, c) : a_entry(Ag, p.a_f32, p.kp, c, lane);
      sA[c * LD + lane] = v;
    }
  }
  if (lane < KMAX) sb[lane] = act_r ? bg[lane] : 0.0;
  __syncthreads();

  // ---- 1. tridiagonalisation ----
  for (int j = 0; j + 2 < k; ++j) {
    const bool act = act_r && lane > j;
    const double x = act ? sA[j * LD + lane] : 0.0;
    const double x0 = sA[j * LD + j + 1];
    const double rest = wave_sum((act && lane > j + 1) ? x * x : 0.0);
    if (rest == 0.0) {
      if (lane == 0) { se[j] = x0; stau[j] = 0.0; }
      __syncthreads();
      continue;
    }
    const double nrm = sqrt(rest + x0 * x0);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;
    const double v0 = x0 - alpha;
    const double tau = 2.0 / (rest + v0 * v0);
    const double v = act ? ((lane == j + 1) ? v0 : x) : 0.0;
    if (lane < KMAX) sv[lane] = v;
    __syncthreads();
    double pr = 0.0;
    if (act)
      for (int c = j + 1; c < k; ++c) pr += sA[c * LD + lane] * sv[c];
    pr *= tau;
    const double K = 0.5 * tau * wave_sum(v * pr);
    const double wr = act ? pr - K * v : 0.0;
    if (lane < KMAX) sw[lane] = wr;
    __syncthreads();
    if (act) {
      for (int c = j + 1; c < k; ++c) sA[c * LD + lane] -= v * sw[c] + wr * sv[c];
      sA[j * LD + lane] = v;   // keep the reflector below the diagonal
    }
    if (lane == 0) { se[j] = alpha; stau[j] = tau; }
    __syncthreads();
  }
  if (act_r) sd[lane] = sA[lane * LD + lane];
  if (lane == 0) {
    if (k >= 2) se[k - 2] = sA[(k - 2) * LD + k - 1];
    se[k - 1] = 0.0;
  }
  __syncthreads();

  // ---- 2. Z = Q = H_0 ... H_{k-3} ----
  for (int c = 0; c < KMAX; ++c)
    if (lane < KMAX) sZ[c * LD + lane] = (lane == c) ? 1.0 : 0.0;
  __syncthreads();
  for (int j = k - 3; j >= 0; --j) {
    const double tau = stau[j];
    if (tau == 0.0) continue;
    const int c = lane;
    if (c > j && c < k) {
      double s = 0.0;
      for (int i = j + 1; i < k; ++i) s += sA[j * LD + i] * sZ[c * LD + i];
      s *= tau;
      for (int i = j + 1; i < k; ++i) sZ[c * LD + i] -= s * sA[j * LD + i];
    }
    __syncthreads();
  }

  // ---- 3. implicit QL on (d, e), rotations applied to Z ----
  const double eps = 2.220446049250313e-16;
  // negligible subdiagonal: relative to the neighbours or to the norm of T
  const double tnorm = wave_max(act_r ? fabs(sd[lane]) + fabs(se[lane]) : 0.0);
  int status = 0;
  for (int l = 0; l < k && !status; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < k - 1; ++m) {
        const double dd = fabs(sd[m]) + fabs(sd[m + 1]);
        if (fabs(se[m]) <= eps * dd || fabs(se[m]) <= eps * tnorm) break;
      }
      if (m == l) break;
      if (++iter > 30) { status = l + 1; break; }
      double g = (sd[l + 1] - sd[l]) / (2.0 * se[l]);
      double r = hypot(g, 1.0);
      g = sd[m] - sd[l] + se[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      double carry = act_r ? sZ[m * LD + lane] : 0.0;
      double ei = se[m - 1], di = sd[m - 1], di1 = sd[m];
      int i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        const double f = s * ei, bb = c * ei;
        r = hypot(f, g);
        se[i + 1] = r;
        if (r == 0.0) {
          sd[i + 1] = di1 - pp;
          se[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = di1 - pp;
        r = (di - g) * s + 2.0 * c * bb;
        pp = s * r;
        sd[i + 1] = g + pp;
        g = c * r - bb;
        const double zi = act_r ? sZ[i * LD + lane] : 0.0;
        if (act_r) sZ[(i + 1) * LD + lane] = s * zi + c * carry;
        carry = c * zi - s * carry;
        if (i > l) {
          ei = se[i - 1];
          di1 = di;
          di = sd[i - 1];
        }
      }
      if (underflow) {
        if (act_r) sZ[(i + 1) * LD + lane] = carry;
        __syncthreads();
        continue;
      }
      if (act_r) sZ[l * LD + lane] = carry;
      sd[l] = sd[l] - pp;
      se[l] = g;
      se[m] = 0.0;
      __syncthreads();
    }
  }
  __syncthreads();

  // ---- 4. eigenvalues of A + (k-1)/rho I ----
  double lam = act_r ? sd[lane] : 0.0;
  if (p.guard) {
    const double mn = wave_min(act_r ? lam : 1e300);
    if (mn < 0.0) lam = lam - mn;
  }
  lam = lam * 1.0 + (double)(k - 1) / p.inflation;

  // ---- 5. overline_w = Z diag(1/lam) Z^T b ----
  {
    double t = 0.0;
    if (act_r)
      for (int i = 0; i < k; ++i) t += sZ[lane * LD + i] * sb[i];
    if (lane < KMAX) st[lane] = act_r ? t / lam : 0.0;
    if (lane < KMAX) sx[lane] = act_r ? sqrt((double)(k - 1) / lam) : 0.0;
  }
  __syncthreads();
  double wbar = 0.0;
  if (act_r)
    for (int j = 0; j < k; ++j) wbar += sZ[j * LD + lane] * st[j];
  // T row r: T(r,c) = sum_j z(r,j) f_j z(c,j); trans(c, r) = T(c, r) + wbar_c
  // lane e keeps column e of trans: tcol[f] = T(f,e) + wbar_f = T(e,f) + wbar_f
  if (lane < KMAX) sw[lane] = act_r ? wbar : 0.0;
  __syncthreads();
  double tcol[KMAX];
#pragma unroll
  for (int f = 0; f < KMAX; ++f) tcol[f] = 0.0;
  if (act_r) {
    double zf[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) zf[j] = (j < k) ? sZ[j * LD + lane] * sx[j] : 0.0;
#pragma unroll
    for (int f = 0; f < KMAX; ++f) {
      if (f < k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
          if (j < k) s += zf[j] * sZ[j * LD + f];
        tcol[f] = s + sw[f];
      }
    }
  }
  // a non-finite overline_w (NaN or Inf in d or R^-1): trans = I, so the
  // update returns the background up to its re-centring
  if (__builtin_amdgcn_ballot_w64(act_r && !__builtin_isfinite(wbar))) {
    status = kStatusNonFinite;
#pragma unroll
    for (int f = 0; f < KMAX; ++f) tcol[f] = (f == lane) ? 1.0 : 0.0;
    wbar = __longlong_as_double(0x7ff8000000000000LL);
  }
  if (p.status) p.status[a] = status;
  if (p.trans_out && act_r) {
    // trans(f, e) column-major: column e is tcol of lane e
    double* tr = p.trans_out + a * (long long)k * k + (long long)lane * k;
#pragma unroll
    for (int f = 0; f < KMAX; ++f)
      if (f < k) tr[f] = tcol[f];
  }
  if (p.w_out && act_r) p.w_out[a * (long long)k + lane] = wbar;
  if (!p.xb) return;

  // ---- 6. update ----
  const int ah = (int)(a / p.nzc), kc = (int)(a % p.nzc);
  const int ic = ah % p.g.nlat_c, jc = ah / p.g.nlat_c;
  int sj, ej, si, ei;
  cube_span(p.g, ic, p.g.nlat, &sj, &ej);
  cube_span(p.g, jc, p.g.nlon, &si, &ei);
  const long long lat_ext = p.g.nlat + 2 * p.halo;
  int sk = 1, ek = p.nz;
  if (p.vertical) {
    sk = kc * p.cube_nz + 1;
    ek = min(sk + p.cube_nz - 1, p.nz);
  }
  for (int lo = si; lo <= ei; ++lo) {
    for (int la = sj; la <= ej; ++la) {
      const long long off = ((long long)(lo - 1 + p.halo) * lat_ext + (la - 1 + p.halo)) * p.n;
      for (int vv = 0; vv < p.n_var; ++vv) {
        const int start = p.var_start[vv] - 1;
        int l0, l1;
        if (p.var_lev[vv] == 1) {
          if (p.vertical && ek != p.nz) continue;
          l0 = start; l1 = start;
        } else if (p.vertical) {
          l0 = start + sk - 1; l1 = start + ek - 1;
        } else {
          l0 = start; l1 = start + p.var_lev[vv] - 1;
        }
        for (int nk = l0; nk <= l1; ++nk) {
          const long long o = off + nk;
          __syncthreads();
          if (act_r) sx[lane] = p.xb[o * k + lane];
          __syncthreads();
          double s = 0.0;
          if (act_r) {
#pragma unroll
            for (int f = 0; f < KMAX; ++f)
              if (f < k) s += tcol[f] * sx[f];
          }
          const double ab = p.axb[o];
          const double xe = act_r ? s + ab : 0.0;
          const double mean = wave_sum(xe) / k;
          if (act_r) p.xa[o * k + lane] = xe - mean;
          if (lane == 0) p.axa[o] = mean;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// caller epilogue (module_da_parallel.F90:219-261)
// ---------------------------------------------------------------------------
__global__ void k_epilogue(EpiParams p) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long lat_ext = p.nlat + 2 * p.halo, lon_ext = p.nlon + 2 * p.halo;
  const long long total = lat_ext * lon_ext * p.n;
  if (idx >= total) return;
  const long long v = idx % p.n;
  const long long la = (idx / p.n) % lat_ext;
  const long long lo = idx / (p.n * lat_ext);
  const int k = p.k;
  const double* b = p.xb + idx * k;
  const double* xa = p.xa + idx * k;
  double* o = p.xa_out + idx * k;
  const bool interior = la >= p.halo && la < p.nlat + p.halo && lo >= p.halo && lo < p.nlon + p.halo;
  if (!interior) {
    for (int e = 0; e < k; ++e) o[e] = 0.0;
    return;
  }
  const long long so = ((lo - p.halo) * p.nlat + (la - p.halo)) * p.n + v;
  const double ab = p.axb[idx], aa = p.axa[idx];
  double sb = 0.0, sa = 0.0;
  if (fabs((double)p.glat[la - p.halo]) <= p.polar_lat) {
    for (int e = 0; e < k; ++e) {
#pragma clang fp contract(off)
      const double out = p.gamma1 * (ab + b[e]) + p.gamma2 * (aa + xa[e]);
      o[e] = out;
      sb += b[e] * b[e];
      sa += (out - aa) * (out - aa);
    }
  } else {
    for (int e = 0; e < k; ++e) {
      o[e] = b[e] + ab;
      sb += b[e] * b[e];
      sa += b[e] * b[e];
    }
  }
  p.spread_b[so] = sqrt(sb / (double)(float)k);
  p.spread_a[so] = sqrt(sa / (double)(float)k);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_site_keys(hipStream_t st, int n_sites, const double* lat, const double* lon,
                            double obs_polar, int nb_lat, int nb_lon, double bdeg_lat,
                            double bdeg_lon, int* key, int* val, int* hist) {
  if (n_sites <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_site_keys, dim3((n_sites + 255) / 256), dim3(256), 0, st, n_sites, lat,
                     lon, obs_polar, nb_lat, nb_lon, bdeg_lat, bdeg_lon, key, val, hist);
  return hipGetLastError();
}

hipError_t launch_site_gather(hipStream_t st, int n_sites, const int* order, const double* lat,
                              const double* lon, const int* row_ptr, double* s_cos, double* s_sin,
                              double* s_lonr, double* s_cosl, double* s_sinl, int* s_row0,
                              int* s_nrow) {
  if (n_sites <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_site_gather, dim3((n_sites + 255) / 256), dim3(256), 0, st, n_sites,
                     order, lat, lon, row_ptr, s_cos, s_sin, s_lonr, s_cosl, s_sinl, s_row0,
                     s_nrow);
  return hipGetLastError();
}

hipError_t launch_select_tile(hipStream_t st, const SelectParams& p, bool fill) {
  const int nt = (p.h1 - p.h0 + p.tile - 1) / p.tile;
  if (nt <= 0) return hipSuccess;
  const int blocks = (nt + 3) / 4;
  if (fill) hipLaunchKernelGGL(k_select_tile<true>, dim3(blocks), dim3(256), 0, st, p);
  else hipLaunchKernelGGL(k_select_tile<false>, dim3(blocks), dim3(256), 0, st, p);
  return hipGetLastError();
}

hipError_t launch_select(hipStream_t st, const SelectParams& p, bool fill) {
  const int ncol = p.h1 - p.h0;
  if (ncol <= 0) return hipSuccess;
  const int blocks = ((ncol + 3) / 4 + 7) / 8 * 8;     // a multiple of 8 (XCDs)
  if (fill) hipLaunchKernelGGL(k_select<true>, dim3(blocks), dim3(256), 0, st, p);
  else hipLaunchKernelGGL(k_select<false>, dim3(blocks), dim3(256), 0, st, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// exclusive prefix sum of the row counts (three kernels, no host round trip:
// chunk scans, a scan of the chunk totals, the chunk offsets added)
// ---------------------------------------------------------------------------
constexpr int kScanChunk = 1024;        // elements per workgroup (256 threads x 4)

__device__ __forceinline__ long long block_excl_scan_256(long long v, long long* sh,
                                                         long long& total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  long long before = 0;
  for (int i = 0; i < w; ++i) before += sh[i];
  total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return before + x - v;
}

__global__ __launch_bounds__(256) void k_scan_chunks(const long long* __restrict__ in,
                                                     long long* __restrict__ out, long long n,
                                                     long long* __restrict__ chunk_tot) {
  __shared__ long long sh[4];
  const long long base = (long long)blockIdx.x * kScanChunk + 4 * threadIdx.x;
  long long v[4], s = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    v[u] = base + u < n ? in[base + u] : 0;
    s += v[u];
  }
  long long total;
  long long ex = block_excl_scan_256(s, sh, total);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (base + u < n) out[base + u] = ex;
    ex += v[u];
  }
  if (threadIdx.x == 0) chunk_tot[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the chunk totals in place
__global__ __launch_bounds__(256) void k_scan_totals(long long* __restrict__ tot, long long nb) {
  __shared__ long long sh[4];
  long long carry = 0;
  for (long long c0 = 0; c0 < nb; c0 += 256) {
    const long long i = c0 + threadIdx.x;
    const long long v = i < nb ? tot[i] : 0;
    long long total;
    const long long ex = block_excl_scan_256(v, sh, total);
    if (i < nb) tot[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_scan_add(long long* __restrict__ out, long long n,
                                                  const long long* __restrict__ chunk_off) {
  const long long base = (long long)blockIdx.x * kScanChunk + 4 * threadIdx.x;
  const long long add = chunk_off[blockIdx.x];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (base + u < n) out[base + u] += add;
}

long long scan_tmp_elems(long long n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

hipError_t launch_exclusive_scan(hipStream_t st, const long long* in, long long* out, long long n,
                                 long long* tmp) {
  if (n <= 0) return hipSuccess;
  const long long nb = (n + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)nb), dim3(256), 0, st, in, out, n, tmp);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(256), 0, st, tmp, nb);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(256), 0, st, out, n,
                     (const long long*)tmp);
  return hipGetLastError();
}

__global__ void k_list_counts(const long long* __restrict__ off, int n_list, int* __restrict__ counts,
                              long long* __restrict__ total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_list && counts) counts[i] = (int)(off[i + 1] - off[i]);
  if (i == 0) *total = off[n_list];
}

hipError_t launch_list_counts(hipStream_t st, const long long* off, int n_list, int* counts,
                              long long* total) {
  hipLaunchKernelGGL(k_list_counts, dim3((unsigned)((n_list + 255) / 256 + (n_list == 0))),
                     dim3(256), 0, st, off, n_list, counts, total);
  return hipGetLastError();
}

__global__ void k_zero_ll(long long* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t launch_zero_ll(hipStream_t st, long long* p, long long n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_ll, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Length buckets of a batch's columns over their local obs count p: a stable
// counting sort (columns keep their order within a bucket, so neighbouring
// columns, which share obs rows in L2, stay close in the queue), buckets
// longest first.  With band_w > 0 the columns enter the sort in bands of
// band_w longitudes, latitude by latitude, so the analyses that one XCD runs
// at a time form a compact patch rather than a run along meridians.  Three
// launches: per-block counts, one workgroup scanning the block counts per
// bucket, and the scatter.

__device__ __forceinline__ int bucket_column(const BucketParams& p, int e) {
  return p.band_w > 0 ? (int)band_index(e, p.n_cols, p.band_w, p.band_l) : e;
}

long long bucket_blk_elems(int n_cols) {
  return (long long)((n_cols + kBucketBlock - 1) / kBucketBlock) * 3 * kBuckets;
}

__global__ __launch_bounds__(256) void k_bucket_blocks(BucketParams p) {
  __shared__ int s_cnt[kBuckets];
  __shared__ unsigned long long s_rows[kBuckets];
  if (threadIdx.x < kBuckets) {
    s_cnt[threadIdx.x] = 0;
    s_rows[threadIdx.x] = 0;
  }
  __syncthreads();
  const int base = blockIdx.x * kBucketBlock + 4 * threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (base + u < p.n_cols) {
      const int r = p.counts[bucket_column(p, base + u)];
      const int b = bucket_of(r);
      atomicAdd(&s_cnt[b], 1);
      atomicAdd(&s_rows[b], (unsigned long long)r);
    }
  }
  __syncthreads();
  if (threadIdx.x < kBuckets) {
    long long* blk = p.blk + (long long)blockIdx.x * 3 * kBuckets;
    blk[threadIdx.x] = s_cnt[threadIdx.x];
    blk[kBuckets + threadIdx.x] = (long long)s_rows[threadIdx.x];
  }
}

// one workgroup: bucket totals, the segment starts (analyses, longest bucket
// first) and each block's first column position within its bucket; zeroes
// the queue counters
__global__ __launch_bounds__(64) void k_bucket_scan(BucketParams p, int n_blk) {
  __shared__ long long s_tot[kBuckets];
  __shared__ long long s_start[kBuckets];
  const int b = threadIdx.x;
  if (b < kBuckets) {
    long long t = 0, r = 0;
    for (int i = 0; i < n_blk; ++i) {
      t += p.blk[(long long)i * 3 * kBuckets + b];
      r += p.blk[(long long)i * 3 * kBuckets + kBuckets + b];
    }
    s_tot[b] = t;
    p.bucket_stats[b] = t * p.nzc;
    p.bucket_stats[kBuckets + b] = r * p.nzc;
    for (int x = 0; x < kQueueParts; ++x) p.q_counters[b * kQueueParts + x] = 0;
    if (p.ticks) {
      p.ticks[b] = 0;
      p.ticks[kBuckets + 4 * b] = ~0ull;
      p.ticks[kBuckets + 4 * b + 1] = 0;
      p.ticks[kBuckets + 4 * b + 2] = 0;
      p.ticks[kBuckets + 4 * b + 3] = 0;
    }
  }
  __syncthreads();
  if (b == 0) {
    long long at = 0;
    for (int j = 0; j < kBuckets; ++j) {
      const int bb = kBuckets - 1 - j;
      s_start[bb] = at;
      p.seg_off[j] = at * p.nzc;
      at += s_tot[bb];
    }
    p.seg_off[kBuckets] = at * p.nzc;
  }
  __syncthreads();
  if (b < kBuckets) {
    long long at = s_start[b];
    for (int i = 0; i < n_blk; ++i) {
      long long* blk = p.blk + (long long)i * 3 * kBuckets;
      blk[2 * kBuckets + b] = at;
      at += blk[b];
    }
  }
}

__global__ __launch_bounds__(256) void k_bucket_scatter(BucketParams p) {
  __shared__ long long sh[4];
  const int base = blockIdx.x * kBucketBlock + 4 * threadIdx.x;
  int bk[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    bk[u] = base + u < p.n_cols ? bucket_of(p.counts[bucket_column(p, base + u)]) : -1;
  const long long* blk = p.blk + (long long)blockIdx.x * 3 * kBuckets;
  for (int b = 0; b < kBuckets; ++b) {
    long long mine = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) mine += bk[u] == b;
    long long total;
    long long rank = block_excl_scan_256(mine, sh, total);
    if (total == 0) continue;
    const long long first = blk[2 * kBuckets + b];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (bk[u] != b) continue;
      const long long pos = first + rank++;
      const int c = bucket_column(p, base + u);
      for (int kc = 0; kc < p.nzc; ++kc) p.order[pos * p.nzc + kc] = c * p.nzc + kc;
    }
  }
}

hipError_t launch_buckets(hipStream_t st, const BucketParams& p) {
  const int n_blk = (p.n_cols + kBucketBlock - 1) / kBucketBlock;
  if (n_blk > 0)
    hipLaunchKernelGGL(k_bucket_blocks, dim3((unsigned)n_blk), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(64), 0, st, p, n_blk);
  if (n_blk > 0)
    hipLaunchKernelGGL(k_bucket_scatter, dim3((unsigned)n_blk), dim3(256), 0, st, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile-union products (horizontal loop, k <= 64).  A workgroup takes a tile
// of NA consecutive analyses; the union of their candidate rows (k_select_tile)
// is staged through LDS CH rows at a time, once for the whole tile, and every
// wavefront keeps its own analysis' rows: it evaluates the same GC weight as
// k_select for each staged row and compacts the rows with weight > 0 (ballot
// + prefix count) into a per-wavefront list, in union (bin) order.  The MFMA
// accumulation is that of the per-analysis kernels, over the listed rows.
// ---------------------------------------------------------------------------
template <int KP, int PREC>
struct TileTypes {
  using Y = typename std::conditional<PREC == 0, double, float>::type;
};

template <int KP, int PREC, int NA, int CH>
__global__ __launch_bounds__(NA * 64) void k_products_tile(ProdParams p) {
  using Y = typename TileTypes<KP, PREC>::Y;
  constexpr int T = KP / 16;
  using PT = PairTable<T>;
  constexpr int NP = PT::N;
  __shared__ __align__(16) Y sY[CH * KP];
  __shared__ int sRow[CH], sSite[CH];
  __shared__ int sLi[NA][CH];
  __shared__ double sW[NA][CH], sWd[NA][CH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, q = lane >> 4;
  const int tile = blockIdx.x;
  const long long ta0 = (long long)tile * p.tile;          // batch-local first analysis
  const long long a_loc = ta0 + wave;
  const bool act = wave < p.tile && a_loc < p.n_an;
  const long long u0 = min(p.row_off[tile], p.row_cap), u1 = min(p.row_off[tile + 1], p.row_cap);
  // this wavefront's analysis centre
  double c2 = 0.0, s2 = 0.0, l2 = 0.0;
  if (act) {
    int sj, ej, si, ei;
    column_center(p.g, p.h0 + (int)a_loc, &sj, &ej, &si, &ei);
    const double clat = (double)p.g.glat[(sj + ej) / 2 - 1];
    const double clon = (double)p.g.glon[(si + ei) / 2 - 1];
    c2 = cos(deg2rad_ref(clat));
    s2 = sin(deg2rad_ref(clat));
    l2 = deg2rad_ref(clon);
  }
  using Acc = typename std::conditional<PREC == 0, d4, f4>::type;
  Acc acc[NP];
#pragma unroll
  for (int x = 0; x < NP; ++x)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[x][r] = 0;
  double bacc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) bacc[t] = 0.0;
  int nsel = 0;
  const int k = p.k;
  for (long long c0 = u0; c0 < u1; c0 += CH) {
    const int nch = (int)min((long long)CH, u1 - c0);
    __syncthreads();
    // ---- stage the chunk's rows (y in the storage precision of the mode) ----
    for (int i = threadIdx.x; i < CH; i += NA * 64) {
      sRow[i] = (i < nch) ? p.row_idx[c0 + i] : 0;
      sSite[i] = (i < nch) ? p.row_site[c0 + i] : 0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < CH * KP; e += NA * 64) {
      const int r = e / KP, m = e % KP;
      Y v = 0;
      if (r < nch && m < k) {
        if constexpr (PREC == 0) v = p.yb64[(long long)sRow[r] * k + m];
        else v = p.yb32[(long long)sRow[r] * k + m];
      }
      sY[e] = v;
    }
    // ---- this analysis' rows of the chunk: the weight of k_select / row_weight ----
    int cnt = 0;
    if (act) {
      double w = 0.0, wd = 0.0;
      bool sel = false;
      if (lane < nch) {
        const int st = sSite[lane];
        const double d = latlon_dist_pre(p.s_cos[st], p.s_sin[st], p.s_lonr[st], c2, s2, l2);
        const double gc = gc_weight(d * 1.0 / p.local_dist, p.gc_filter);
        if (gc > 0.0) {
          const int row = sRow[lane];
          w = row_weight(p, row, gc, 0.0);
          wd = w * p.dep[row];
          sel = true;
        }
      }
      const unsigned long long m = __ballot(sel);
      cnt = __popcll(m);
      if (sel) {
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        sLi[wave][rank] = lane;
        sW[wave][rank] = w;
        sWd[wave][rank] = wd;
      }
    }
    __syncthreads();
    if (!act) continue;
    nsel += cnt;
    if constexpr (PREC == 2) {
      // BF16 hi/lo split of sqrt(w) y, 32 rows per MFMA (8 per lane)
      for (int g = 0; g < cnt; g += 32) {
        bf8 hi[T], lo[T];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int idx = g + 8 * q + u;
          float sq = 0.f;
          double wd = 0.0;
          int li = 0;
          if (idx < cnt) {
            li = sLi[wave][idx];
            sq = sqrtf((float)sW[wave][idx]);
            wd = sWd[wave][idx];
          }
#pragma unroll
          for (int t = 0; t < T; ++t) {
            const float yv = (idx < cnt) ? (float)sY[li * KP + 16 * t + i16] : 0.f;
            const float yh = sq * yv;
            const __bf16 h = (__bf16)yh;
            hi[t][u] = h;
            lo[t][u] = (__bf16)(yh - (float)h);
            bacc[t] = fma(wd, (double)yv, bacc[t]);
          }
        }
#pragma unroll
        for (int x = 0; x < NP; ++x) {
          const int I = PT::I(x), J = PT::J(x);
          acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[I], hi[J], acc[x], 0, 0, 0);
          acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[I], lo[J], acc[x], 0, 0, 0);
          acc[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[I], hi[J], acc[x], 0, 0, 0);
        }
      }
    } else {
      // FP64 / FP32: 4 rows per MFMA (one per lane quarter)
      for (int g = 0; g < cnt; g += 4) {
        const int idx = g + q;
        Y y[T];
        double w = 0.0, wd = 0.0;
        if (idx < cnt) {
          const int li = sLi[wave][idx];
          w = sW[wave][idx];
          wd = sWd[wave][idx];
#pragma unroll
          for (int t = 0; t < T; ++t) y[t] = sY[li * KP + 16 * t + i16];
        } else {
#pragma unroll
          for (int t = 0; t < T; ++t) y[t] = 0;
        }
#pragma unroll
        for (int x = 0; x < NP; ++x) {
          const int I = PT::I(x), J = PT::J(x);
          if constexpr (PREC == 0)
            acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[I], w * y[J], acc[x], 0, 0, 0);
          else
            acc[x] = __builtin_amdgcn_mfma_f32_16x16x4f32(y[I], (float)w * y[J], acc[x], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < T; ++t) bacc[t] = fma(wd, (double)y[t], bacc[t]);
      }
    }
  }
  if (!act) return;
  const long long al = a_loc;
  double* A = p.A + al * (long long)KP * KP;
#pragma unroll
  for (int x = 0; x < NP; ++x) {
    const int I = PT::I(x), J = PT::J(x);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // f64 16x16x4 output: D[l/16 + 4r][l%16]; f32 16x16x4 / bf16 16x16x32: D[4(l/16) + r][l%16]
      const int row = (PREC == 0) ? 16 * I + q + 4 * r : 16 * I + 4 * q + r;
      const int col = 16 * J + i16;
      const double v = (double)acc[x][r];
      A[(long long)col * KP + row] = v;
      if (I != J) A[(long long)row * KP + col] = v;
    }
  }
  double* b = p.b + al * KP;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    double v = bacc[t];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (q == 0) b[16 * t + i16] = v;
  }
  if (lane == 0) p.nobsl[al] = nsel;
}

hipError_t launch_products_tile(hipStream_t st, const ProdParams& p, int precision) {
  constexpr int NA = 8, CH = 64;
  if (p.tile != NA) return hipErrorInvalidValue;
  const long long nt = (p.n_an + NA - 1) / NA;
  if (nt <= 0) return hipSuccess;
  const dim3 grid((unsigned)nt), block(NA * 64);
  if (p.kp == 32) {
    if (precision == 0) hipLaunchKernelGGL((k_products_tile<32, 0, NA, CH>), grid, block, 0, st, p);
    else if (precision == 1) hipLaunchKernelGGL((k_products_tile<32, 1, NA, CH>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((k_products_tile<32, 2, NA, CH>), grid, block, 0, st, p);
  } else if (p.kp == 64) {
    if (precision == 0) hipLaunchKernelGGL((k_products_tile<64, 0, NA, CH>), grid, block, 0, st, p);
    else if (precision == 1) hipLaunchKernelGGL((k_products_tile<64, 1, NA, CH>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((k_products_tile<64, 2, NA, CH>), grid, block, 0, st, p);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Workgroups of kernel f resident on the whole device at once (the grid of a
// queue-schedule launch), computed once per kernel.
static int resident_blocks(const void* f, int threads, std::atomic<int>& cache) {
  int v = cache.load(std::memory_order_relaxed);
  if (v > 0) return v;
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, threads, 0) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  v = std::max(1, per_cu) * std::max(1, cus);
  cache.store(v, std::memory_order_relaxed);
  return v;
}

// One workgroup per analysis: its wavefronts (one, or four sharing the
// analysis) start and retire together, so no wavefront waits on another
// analysis of its workgroup.
template <class K>
static hipError_t launch_prod(hipStream_t st, const ProdParams& p, K* kern, long long blocks,
                              std::atomic<int>& cache, int wpa, long long* capacity) {
  ProdParams q = p;
  unsigned grid;
  constexpr int apb = 1;
  const int threads = 64 * wpa;
  const int r = (q.order || capacity)
                    ? resident_blocks(reinterpret_cast<const void*>(kern), threads, cache)
                    : 1;
  if (r <= 0) return hipErrorInvalidDeviceFunction;
  if (q.q_counter && q.order) {
    grid = (unsigned)std::min<long long>(r, blocks);
    q.n_blocks = (int)grid;
  } else if (q.order) {
    // an XCD's share is an eighth of each bucket, rounded: at most
    // n / 8 + kBuckets analyses
    const long long per_xcd = ((q.n_an + 7) / 8 + kBuckets + apb - 1) / apb;
    grid = (unsigned)(8 * per_xcd);
    q.n_blocks = (int)grid;
  } else {
    q.n_blocks = (int)blocks;
    grid = (unsigned)(((blocks + 7) / 8) * 8);
  }
  // wavefront groups the device holds at once (the load balance's capacity)
  if (capacity) *capacity = (long long)std::min<long long>(r, grid) * apb;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, st, q);
  return hipGetLastError();
}

template <int KP, int SCHED>
static hipError_t launch_products_sched(hipStream_t st, const ProdParams& p, int precision,
                                        long long blocks, long long* capacity) {
  constexpr int T = KP / 16;
  constexpr int NPAIR = T * (T + 1) / 2;
  constexpr int WPA = (NPAIR + 9) / 10;
  // PW: the row lists carry the weights w and w d (horizontal loop), so the
  // products read them instead of forming them per row; straight-line code
  // lets a group's loads be in flight together
  static std::atomic<int> c64[2], c32[2], c16a[2], c16b[2];
  const int w = p.pre_weighted ? 1 : 0;
  if (precision == 0)
    return w ? launch_prod(st, p, k_products_f64<KP, SCHED, true>, blocks, c64[1], WPA, capacity)
             : launch_prod(st, p, k_products_f64<KP, SCHED, false>, blocks, c64[0], WPA, capacity);
  if (precision == 1)
    return w ? launch_prod(st, p, k_products_f32<KP, SCHED, true>, blocks, c32[1], WPA, capacity)
             : launch_prod(st, p, k_products_f32<KP, SCHED, false>, blocks, c32[0], WPA, capacity);
  if (WPA == 1 && p.groups == 2)
    return w ? launch_prod(st, p, k_products_bf16<KP, 2, SCHED, true>, blocks, c16b[1], WPA,
                           capacity)
             : launch_prod(st, p, k_products_bf16<KP, 2, SCHED, false>, blocks, c16b[0], WPA,
                           capacity);
  return w ? launch_prod(st, p, k_products_bf16<KP, 1, SCHED, true>, blocks, c16a[1], WPA,
                         capacity)
           : launch_prod(st, p, k_products_bf16<KP, 1, SCHED, false>, blocks, c16a[0], WPA,
                         capacity);
}

template <int KP>
static hipError_t launch_products_kp(hipStream_t st, const ProdParams& p, int precision,
                                     long long* capacity) {
  const long long blocks = p.n_an;
  if (blocks <= 0) return hipSuccess;
  // order set: a bucket segment, in dispatch order or (q_counter set) on a
  // persistent queue; the ticks select the timed instantiation
  const bool timed = p.bucket_ticks != nullptr;
  if (!p.order) return launch_products_sched<KP, 0>(st, p, precision, blocks, capacity);
  if (!p.q_counter)
    return timed ? launch_products_sched<KP, 2>(st, p, precision, blocks, capacity)
                 : launch_products_sched<KP, 1>(st, p, precision, blocks, capacity);
  return timed ? launch_products_sched<KP, 4>(st, p, precision, blocks, capacity)
               : launch_products_sched<KP, 3>(st, p, precision, blocks, capacity);
}

hipError_t launch_products_queue(hipStream_t st, const ProdParams& p, int precision,
                                 long long* capacity) {
  if (!p.order || !p.seg_off || p.seg_lo >= p.seg_hi) return hipErrorInvalidValue;
  if (p.kp == 32) return launch_products_kp<32>(st, p, precision, capacity);
  if (p.kp == 64) return launch_products_kp<64>(st, p, precision, capacity);
  if (p.kp == 128) return launch_products_kp<128>(st, p, precision, capacity);
  return hipErrorInvalidValue;
}

hipError_t launch_products(hipStream_t st, const ProdParams& p, int precision) {
  if (p.kp == 32) return launch_products_kp<32>(st, p, precision, nullptr);
  if (p.kp == 64) return launch_products_kp<64>(st, p, precision, nullptr);
  if (p.kp == 128) return launch_products_kp<128>(st, p, precision, nullptr);
  return hipErrorInvalidValue;
}

hipError_t launch_transform(hipStream_t st, const XformParams& p) {
  if (p.n_an <= 0) return hipSuccess;
  // with a list, n_an is the list length
  if (p.k <= 32) hipLaunchKernelGGL(k_transform<32>, dim3((unsigned)p.n_an), dim3(64), 0, st, p);
  else if (p.k <= 64) hipLaunchKernelGGL(k_transform<64>, dim3((unsigned)p.n_an), dim3(64), 0, st, p);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// A wavefront per state element: lane e holds members e and e + 64.
template <int R>
__global__ __launch_bounds__(256) void k_split(SplitParams p) {
  const int lane = threadIdx.x & 63;
  const long long el = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (el >= p.n_elems) return;
  const double* x = p.x_in + el * p.k;
  double v[R], sum = 0.0;
  bool bad = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = lane + 64 * r;
    v[r] = e < p.k ? x[e] : 0.0;
    bad = bad || (v[r] != v[r]);
    sum += v[r];
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, 64);
  const double mean = sum / p.k;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = lane + 64 * r;
    if (e >= p.k) continue;
    if (p.xb) p.xb[el * p.k + e] = v[r] - mean;
    else p.xb32[el * p.k + e] = (float)(v[r] - mean);
  }
  if (lane == 0) p.axb[el] = mean;
  if (__any(bad) && lane == 0) {
    atomicAdd(&p.nan[0], 1ull);
    atomicMax(&p.nan[1], (unsigned long long)(p.n_elems - el));
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

hipError_t launch_split(hipStream_t st, const SplitParams& p) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p.nan, 2);
  if (p.n_elems <= 0) return hipGetLastError();
  const dim3 grid((unsigned)((p.n_elems + 3) / 4));
  if (p.k <= 64) hipLaunchKernelGGL(k_split<1>, grid, dim3(256), 0, st, p);
  else if (p.k <= 128) hipLaunchKernelGGL(k_split<2>, grid, dim3(256), 0, st, p);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// One workgroup per column (halo columns included: they are zeroed, as
// letkf_compute.F90:191-193 zeroes xa and axa before the loop); wavefront w
// takes the state elements v = w, w + 4, ...  For each covering cube in loop
// order (lon cubes outer, lat cubes inner), lane e forms
// dist(e) = axb(v) + sum_m xb(m, v) trans(m, e) (xb(m, v) broadcast from lane
// m), the wavefront sums dist for the mean, and the lane adds dist(e) - mean
// to xa(e, v) and the mean to axa(v) (xb_add2_xa, letkf_compute.F90:332-382);
// the column is then divided by its cube count (:248-255).
template <int KM>
__global__ __launch_bounds__(256) void k_overlap_update(OverlapParams p) {
  constexpr int R = KM / 64;            // members per lane
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lat_ext = p.g.nlat + 2 * p.halo;
  const int la_e = blockIdx.x % lat_ext, lo_e = blockIdx.x / lat_ext;
  const int k = p.k, n = p.n;
  const long long off = ((long long)lo_e * lat_ext + la_e) * n;
  const int la = la_e - p.halo, lo = lo_e - p.halo;
  if (la < 0 || la >= p.g.nlat || lo < 0 || lo >= p.g.nlon) {
    for (int i = threadIdx.x; i < n * k; i += blockDim.x) p.xa[off * k + i] = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) p.axa[off + i] = 0.0;
    return;
  }
  const int cn = p.g.cube_n, cs = p.g.cube_s;
  // cubes c with c cs <= x <= c cs + cn - 1 (0-based x)
  const int ic0 = la - cn + 1 <= 0 ? 0 : (la - cn + cs) / cs;
  const int ic1 = min(la / cs, p.g.nlat_c - 1);
  const int jc0 = lo - cn + 1 <= 0 ? 0 : (lo - cn + cs) / cs;
  const int jc1 = min(lo / cs, p.g.nlon_c - 1);
  const double count = (double)((ic1 - ic0 + 1) * (jc1 - jc0 + 1));
  const long long kk = (long long)k * k;
  for (int v = wave; v < n; v += 4) {
    const double* xbc = p.xb + (off + v) * k;
    const double ab = p.axb[off + v];
    double xv[R], acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = lane + 64 * r;
      xv[r] = e < k ? xbc[e] : 0.0;
      acc[r] = 0.0;
    }
    double accm = 0.0;
    for (int jc = jc0; jc <= jc1; ++jc) {
      for (int ic = ic0; ic <= ic1; ++ic) {
        const long long c = (long long)jc * p.g.nlat_c + ic;
        const bool keep = p.status[c] != 0;
        double dist[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int e = lane + 64 * r;
          const double* tr = p.trans + c * kk + (long long)(e < k ? e : 0) * k;
          double s = 0.0;
          if (keep) {
            s = xv[r];
          } else {
#pragma unroll
            for (int rr = 0; rr < R; ++rr)
              for (int m = 0; m < min(64, k - 64 * rr); ++m)
                s = fma(__shfl(xv[rr], m, 64), tr[64 * rr + m], s);
          }
          dist[r] = e < k ? ab + s : 0.0;
        }
        double sum = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) sum += dist[r];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, 64);
        const double mean = sum / k;
        accm += mean;
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += dist[r] - mean;
      }
    }
    double* xac = p.xa + (off + v) * k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = lane + 64 * r;
      if (e < k) xac[e] = acc[r] / count;
    }
    if (lane == 0) p.axa[off + v] = accm / count;
  }
}

hipError_t launch_overlap_update(hipStream_t st, const OverlapParams& p) {
  const long long cols = (long long)(p.g.nlat + 2 * p.halo) * (p.g.nlon + 2 * p.halo);
  if (cols <= 0) return hipSuccess;
  if (p.k <= 64)
    hipLaunchKernelGGL(k_overlap_update<64>, dim3((unsigned)cols), dim3(256), 0, st, p);
  else if (p.k <= 128)
    hipLaunchKernelGGL(k_overlap_update<128>, dim3((unsigned)cols), dim3(256), 0, st, p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_epilogue(hipStream_t st, const EpiParams& p) {
  const long long total = (long long)(p.nlat + 2 * p.halo) * (p.nlon + 2 * p.halo) * p.n;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_epilogue, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, p);
  return hipGetLastError();
}

}  // namespace dida
