// letkf_eig.hip -- FP64 eigensolve + transform + update in three kernels.
//
// The k x k matrices M = A + (k-1)/rho I of one batch of local analyses are
// diagonalised by Householder tridiagonalisation and implicit QL with
// Wilkinson shift (Golub & Van Loan Alg. 8.3.1, Sec. 8.3.5; Bowdler, Martin,
// Reinsch & Wilkinson 1968), split so that every phase runs with the
// parallelism it has:
//
//   k_tridiag   one wavefront per matrix, lane r keeps row r of A in VGPRs;
//               writes the reflectors, tau, and the tridiagonal (d, e)
//   k_ql_log    one LANE per matrix (64 matrices per wavefront): the scalar
//               QL recurrence, which is a serial chain per matrix; writes the
//               eigenvalues and a log of the Givens rotations of every chase
//   k_vec       one wavefront per matrix (k <= 64), lane r keeps row r of Z
//               in VGPRs: Q = H_0 ... H_{k-3} by forward accumulation, the
//               logged rotations applied in order (V = Q Z), then
//               overline_w = V diag(1/lambda) V^T b and
//               trans = V diag(sqrt((k-1)/lambda)) V^T + overline_w 1^T
//               (letkf_core.F90:129-225) on f64 MFMA
//   k_update    the ensemble update xa = trans^T xb + axb, re-centred
//               (xb_2_xa, letkf_compute.F90:259-329, or the per-level update
//               of the vertical loop, letkf_compute.F90:501-552)
// k = 65..128 runs k_tridiag128, k_ql_log<128>, k_vec128 and k_update128
// with half rows per thread (256 threads per matrix).
//
// Matrices whose QL did not converge or whose rotation log overflowed are
// flagged (status != 0) and appended to a device-side list; k_vec and
// k_update skip them.  A second pass of k_ql_log, k_vec and k_update over
// that list (list mode: a longer rotation log, more QL iterations) redoes
// them from the stored tridiagonal form, with no host synchronisation.  An
// analysis still flagged after it leaves xa = xb, axa = axb for its state and
// NaN in trans/overline_w; the caller gets the count.
#include "letkf_device.h"
#include "letkf_kernels.h"

namespace dida {

// The analysis of workgroup (or lane) `slot`: a = global index, ai = index of
// the batch-local inputs (A, b, refl, tau, d, e), ar = index of the pass's own
// scratch (rot, hdr, nchase, lam).  False when the slot holds no analysis.
__device__ __forceinline__ bool eig_slot(const EigParams& p, long long slot, long long& a,
                                         long long& ai, long long& ar) {
  if (p.list) {
    const int n = min(*p.list_n, p.list_cap);
    if (slot >= n) return false;
    a = p.list[slot];
    ai = a - p.a0;
    ar = slot;
  } else {
    if (slot >= p.n_an) return false;
    a = p.a0 + slot;
    ai = slot;
    ar = slot;
  }
  return true;
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// State rows of one analysis: the columns of its cube and, on the vertical
// loop, the levels of its chunk (with the 2-D variables on the top chunk).
struct AnRows {
  int sj, ej, si, ei, sk, ek;
  long long lat_ext;
};
__device__ __forceinline__ AnRows an_rows(const EigParams& p, long long a) {
  AnRows r;
  const int ah = (int)(a / p.nzc), kc = (int)(a % p.nzc);
  const int ic = ah % p.g.nlat_c, jc = ah / p.g.nlat_c;
  cube_span(p.g, ic, p.g.nlat, &r.sj, &r.ej);
  cube_span(p.g, jc, p.g.nlon, &r.si, &r.ei);
  r.lat_ext = p.g.nlat + 2 * p.halo;
  r.sk = 1;
  r.ek = p.nz;
  if (p.vertical) {
    r.sk = kc * p.cube_nz + 1;
    r.ek = min(r.sk + p.cube_nz - 1, p.nz);
  }
  return r;
}

// trans and overline_w of an analysis whose eigensolve failed on both passes
__device__ void nan_transform(const EigParams& p, long long a) {
  const int k = p.k;
  const double nan = __longlong_as_double(0x7ff8000000000000LL);
  if (p.trans_out)
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) p.trans_out[a * (long long)k * k + e] = nan;
  if (p.w_out)
    for (int e = threadIdx.x; e < k; e += blockDim.x) p.w_out[a * (long long)k + e] = nan;
}

// ---------------------------------------------------------------------------
// k_tridiag
// ---------------------------------------------------------------------------
// arow[c] for a wavefront-uniform column c: a switch lowered to a short tree
// of scalar branches (no per-column selects or dynamic register indexing)
template <int N>
__device__ __forceinline__ double pick_column(const double (&arow)[N], int c) {
  double x = 0.0;
  switch (c) {
#define DIDA_COL_CASE(i) \
  case i:                \
    if constexpr (i < N) x = arow[i]; \
    break;
#define DIDA_COL_CASE8(b) \
  DIDA_COL_CASE(b) DIDA_COL_CASE(b + 1) DIDA_COL_CASE(b + 2) DIDA_COL_CASE(b + 3) \
  DIDA_COL_CASE(b + 4) DIDA_COL_CASE(b + 5) DIDA_COL_CASE(b + 6) DIDA_COL_CASE(b + 7)
    DIDA_COL_CASE8(0) DIDA_COL_CASE8(8) DIDA_COL_CASE8(16) DIDA_COL_CASE8(24)
    DIDA_COL_CASE8(32) DIDA_COL_CASE8(40) DIDA_COL_CASE8(48) DIDA_COL_CASE8(56)
#undef DIDA_COL_CASE8
#undef DIDA_COL_CASE
    default:
      break;
  }
  return x;
}

// k = 64 is capped at 168 registers (three wavefronts per SIMD: the kernel is
// bound by the latency of its per-step reductions and LDS round trips, and a
// few spilled scalars cost less than the third wavefront gains)
template <int KMAX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(KMAX == 64 ? 3 : 2)))
void k_tridiag(EigParams p) {
  constexpr int NB = KMAX / 8;
  __shared__ __align__(16) double sv[KMAX];
  __shared__ __align__(16) double sw[KMAX];
  __shared__ double se_s[KMAX + 1], st_s[KMAX + 1];
  __shared__ double sxn[64];             // column j+1 after step j's update
  const int lane = threadIdx.x;
  const long long a = p.a0 + blockIdx.x;
  if (a >= p.a0 + p.n_an) return;
  const long long al = a - p.a0;
  const int k = p.k, kp = p.kp;
  const double* __restrict__ Ag = p.A + al * (long long)kp * kp;
  double* __restrict__ refl = p.refl + al * (long long)kp * kp;
  const bool in_r = lane < k;

  double arow[KMAX];
  if (p.a_f32) {
    // FP32 in tile order (a_entry; kp = KMAX): one base address per lane and
    // a constant offset per column, every load unconditional and in bounds
    // (lanes past kp read position 0), then a select
    const float* __restrict__ Af =
        reinterpret_cast<const float*>(Ag) + a_tile_pos(lane < KMAX ? lane : 0, KMAX);
    float av[KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c) av[c] = Af[a_tile_pos(c, KMAX) * KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c) arow[c] = (in_r && c < k) ? (double)av[c] : 0.0;
  } else {
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      arow[c] = (in_r && c < k) ? Ag[(long long)c * kp + lane] : 0.0;
  }
  double xcol = arow[0];

  for (int j = 0; j + 2 < k; ++j) {
    const bool act = in_r && lane > j;
    const double x = act ? xcol : 0.0;
    const double x0 = readlane_d(xcol, j + 1);
    const double rest = wave_sum_full((act && lane > j + 1) ? x * x : 0.0);
    __builtin_amdgcn_sched_barrier(0);
    if (rest == 0.0) {
      // column already in tridiagonal form; column j+1 goes through LDS
      sxn[lane] = pick_column(arow, j + 1);
      if (lane < KMAX) refl[(long long)j * kp + lane] = 0.0;
      if (lane == 0) { se_s[j] = x0; st_s[j] = 0.0; }
      __syncthreads();
      xcol = sxn[lane];
      continue;
    }
    const double nrm = sqrt(rest + x0 * x0);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;
    const double v0 = x0 - alpha;
    const double tau = 2.0 / (rest + v0 * v0);
    const double v = act ? ((lane == j + 1) ? v0 : x) : 0.0;
    if (lane < KMAX) sv[lane] = v;
    __syncthreads();
    // p = tau A v over the trailing columns (sv is zero for c <= j); four
    // partial sums, one block of 8 columns in flight at a time
    double pr0 = 0.0, pr1 = 0.0, pr2 = 0.0, pr3 = 0.0;
#pragma unroll
    for (int B = 0; B < NB; ++B) {
      if (8 * B + 7 <= j) continue;
      pr0 = fma(arow[8 * B + 0], sv[8 * B + 0], pr0);
      pr1 = fma(arow[8 * B + 1], sv[8 * B + 1], pr1);
      pr2 = fma(arow[8 * B + 2], sv[8 * B + 2], pr2);
      pr3 = fma(arow[8 * B + 3], sv[8 * B + 3], pr3);
      pr0 = fma(arow[8 * B + 4], sv[8 * B + 4], pr0);
      pr1 = fma(arow[8 * B + 5], sv[8 * B + 5], pr1);
      pr2 = fma(arow[8 * B + 6], sv[8 * B + 6], pr2);
      pr3 = fma(arow[8 * B + 7], sv[8 * B + 7], pr3);
      __builtin_amdgcn_sched_barrier(0);
    }
    double pr = (pr0 + pr1) + (pr2 + pr3);
    pr = act ? pr * tau : 0.0;
    const double K = 0.5 * tau * wave_sum_full(v * pr);
    __builtin_amdgcn_sched_barrier(0);
    const double wr = act ? pr - K * v : 0.0;
    if (lane < KMAX) sw[lane] = wr;
    __syncthreads();
    // rank-2 update A -= v w^T + w v^T, a block of 8 columns per LDS round
    // trip (sv and sw are zero for c <= j); column j+1 goes through LDS after
    // the whole update
#pragma unroll
    for (int B = 0; B < NB; ++B) {
      if (8 * B + 7 <= j) continue;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int c = 8 * B + t;
        arow[c] = fma(-wr, sv[c], fma(-v, sw[c], arow[c]));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    sxn[lane] = pick_column(arow, j + 1);
    if (lane < KMAX) refl[(long long)j * kp + lane] = v;
    if (lane == 0) { se_s[j] = alpha; st_s[j] = tau; }
    __syncthreads();
    xcol = sxn[lane];
  }
  // diagonal and the last subdiagonal entry
  double dr = 0.0, ek2 = 0.0;
#pragma unroll
  for (int c = 0; c < KMAX; ++c) {
    dr = (c == lane) ? arow[c] : dr;
    ek2 = (c == k - 2) ? arow[c] : ek2;
  }
  ek2 = readlane_d(ek2, k - 1);
  __syncthreads();
  double* __restrict__ dg = p.d + al * kp;
  double* __restrict__ eg = p.e + al * kp;
  double* __restrict__ tg = p.tau + al * kp;
  if (lane < KMAX) {
    dg[lane] = in_r ? dr : 0.0;
    double ev = 0.0, tv = 0.0;
    if (lane + 2 < k) { ev = se_s[lane]; tv = st_s[lane]; }
    else if (lane == k - 2) ev = ek2;
    eg[lane] = ev;
    tg[lane] = tv;
  }
}

// ---------------------------------------------------------------------------
// k_ql_log: one lane per matrix
// ---------------------------------------------------------------------------
// d and e of the 64 matrices live in LDS as [i][lane], behind one padding
// row (index -1) so that the sweep's prefetch of i - 1 stays in bounds at
// i = 0.
template <int KMAX>
__global__ __launch_bounds__(64) void k_ql_log(EigParams p) {
  __shared__ double sd_l[KMAX + 1][64];
  __shared__ double se_l[KMAX + 1][64];
  const int lane = threadIdx.x;
#define SD(i) sd_l[(i) + 1][lane]
#define SE(i) se_l[(i) + 1][lane]
  SD(-1) = 0.0;
  SE(-1) = 0.0;
  long long a = 0, ai = 0, ar = 0;
  const bool valid = eig_slot(p, (long long)blockIdx.x * 64 + lane, a, ai, ar);
  const int k = p.k, kp = p.kp;
  double tnorm = 0.0;
  for (int i = 0; i < KMAX; ++i) {
    double dv = 0.0, ev = 0.0;
    if (valid && i < k) {
      dv = p.d[ai * kp + i];
      ev = p.e[ai * kp + i];
    }
    SD(i) = dv;
    SE(i) = ev;
    tnorm = fmax(tnorm, fabs(dv) + fabs(ev));
  }
  if (!valid) return;
  const double eps = 2.220446049250313e-16;
  double2* __restrict__ rot = p.rot + ar * (long long)p.rot_cap;
  int4* __restrict__ hdr = p.hdr + ar * (long long)p.hdr_cap;
  int nrot = 0, nhdr = 0, status = 0;
  for (int l = 0; l < k && status == 0; ++l) {
    int iter = 0;
    for (;;) {
      // first m >= l with a negligible e[m] (a loop with one exit: no
      // per-iteration break bookkeeping)
      int m = l;
      double dm = SD(l);
      bool go = m < k - 1;
      while (go) {
        const double dm1 = SD(m + 1);
        const double em = fabs(SE(m));
        go = !(em <= eps * (fabs(dm) + fabs(dm1)) || em <= eps * tnorm);
        if (go) {
          ++m;
          dm = dm1;
          go = m < k - 1;
        }
      }
      if (m == l) break;
      if (++iter > p.max_iter) { status = l + 1; break; }
      if (nhdr >= p.hdr_cap || nrot + (m - l) + 14 > p.rot_cap) { status = kStatusLogFull; break; }
      double g = (SD(l + 1) - SD(l)) / (2.0 * SE(l));
      double r = hypot(g, 1.0);
      g = SD(m) - SD(l) + SE(l) / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      // the chase's log is block-aligned: positions m .. 8 ceil(m/8) - 1
      // (above it) and 8 floor(lo/8) .. lo - 1 (below) hold the identity, so
      // a reader takes whole blocks of 8 positions without masking
      const int off = nrot;
      for (int q = ((m + 7) & ~7) - 1; q >= m; --q) rot[nrot++] = make_double2(1.0, 0.0);
      int i = m - 1;
      bool underflow = false;
      int iu = l - 1;
      // d[i], e[i] of the next step are read before this step's stores.  The
      // body is straight-line: pointers walk down the lane's LDS column and
      // its log, and an underflow (r2 == 0, rare) is repaired after the
      // step's stores under a wavefront-uniform test, then ends the lane's
      // loop; so the lanes' divergence costs one exit test per step
      double ei = SE(i), di = SD(i), di1 = SD(i + 1);
      double* pd = &SD(i);
      double* pe = &SE(i);
      double2* pr = rot + nrot;
      for (; i >= l; --i, pd -= 64, pe -= 64, ++pr) {
        const double ein = pe[-64], din = pd[-64];
        const double f = s * ei, bb = c * ei;
        // one reciprocal square root instead of sqrt and a divide on the
        // serial chain: r = sqrt(f^2 + g^2), s = f / r, c = g / r
        const double r2 = fma(f, f, g * g);
        const double rinv = rsqrt(r2);
        const double sd_uz = di1 - pp;
        s = f * rinv;
        c = g * rinv;
        g = di1 - pp;
        r = (di - g) * s + 2.0 * c * bb;
        pp = s * r;
        pe[64] = r2 * rinv;
        pd[64] = g + pp;
        g = c * r - bb;
        *pr = make_double2(c, s);
        if (__builtin_amdgcn_ballot_w64(r2 == 0.0)) {
          if (r2 == 0.0) {
            // the chase stops at i: e[i+1] = 0, d[i+1] = d[i+1] - p, no
            // rotation logged for this step
            pe[64] = 0.0;
            pd[64] = sd_uz;
            underflow = true;
            iu = i;
            i = l;
            --pr;
          }
        }
        ei = ein;
        di1 = di;
        di = din;
      }
      nrot = (int)(pr - rot);
      if (underflow) SE(m) = 0.0;
      const int lo = underflow ? iu + 1 : l;
      for (int q = lo - 1; q >= (lo & ~7); --q) rot[nrot++] = make_double2(1.0, 0.0);
      if (m - lo > 0) hdr[nhdr++] = make_int4(lo, m, off, nrot - off);
      else nrot = off;
      if (underflow) continue;
      SD(l) -= pp;
      SE(l) = g;
      SE(m) = 0.0;
    }
  }
  for (int i = 0; i < k; ++i) p.lam[ar * kp + i] = SD(i);
  p.nchase[ar] = nhdr;
  p.status[a] = status;
  if (status != 0 && !p.list) {
    const int s = atomicAdd(p.flag_n, 1);
    if (s < p.flag_cap) p.flag_list[s] = (int)a;
    else p.status[a] = kStatusNotRedone;
    atomicAdd(p.flag_total, 1);
  }
}
#undef SD
#undef SE

// Sum over the NW wavefronts of a workgroup; every thread gets the same value
// (wave sums, then the partials in a fixed order).  `red` holds two slots of
// NW partials so that back-to-back calls need one barrier each.
template <int NW>
__device__ __forceinline__ double block_sum(double v, double* red, int& slot) {
  const double w = wave_sum_full(v);
  if constexpr (NW == 1) {
    return uniform(w);
  } else {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[NW * slot + wv] = w;
    __syncthreads();
    double r = red[NW * slot];
#pragma unroll
    for (int i = 1; i < NW; ++i) r += red[NW * slot + i];
    slot ^= 1;
    return uniform(r);
  }
}

__device__ __forceinline__ double block2_min(double v, double* red, int& slot) {
  const double w = wave_min_full(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[2 * slot + wv] = w;
  __syncthreads();
  const double r = fmin(red[2 * slot], red[2 * slot + 1]);
  slot ^= 1;
  return uniform(r);
}

// ---------------------------------------------------------------------------
// k_vec + k_update: the k_apply work in two kernels sized for two wavefronts
// per SIMD.  k_vec forms V (Q accumulation, logged rotations), overline_w and
// trans = T + overline_w 1^T with LDS staged in small chunks, and writes
// trans in the register layout of the update's MFMA B operand over the
// analysis' reflector block (consumed by then).  k_update streams the state
// vectors of the analysis through that operand.
// ---------------------------------------------------------------------------
// The rotations of positions 8B .. 8B+7 of a chase: the log holds whole
// blocks of 8 positions (identity outside the chase), position i at ring
// entry base - i, so a block is eight consecutive entries read from one
// address (the ring keeps a copy of its first 8 entries past its end, so a
// run never wraps).  base is wavefront-uniform.
template <int RING>
__device__ __forceinline__ void ring_block(const double2* sring, int base, int B,
                                           double2 (&cs)[8]) {
  const double2* e = sring + ((base - 8 * B - 7) & (RING - 1));
#pragma unroll
  for (int t = 0; t < 8; ++t) cs[t] = e[7 - t];
}

// Read-only global memory addressed uniformly across the wavefront: loads
// through these pointer types go to the scalar unit (s_load).  Only for data
// that no kernel writes while the reader runs.
typedef const __attribute__((address_space(4))) int* uniform_int_cptr;
typedef const __attribute__((address_space(4))) double* uniform_dbl_cptr;
// a row of zeros in read-only global memory, for scalar reads past the last
// reflector
__device__ const double kZeroRow[128] = {};
__device__ __forceinline__ int4 load_uniform_int4(uniform_int_cptr q, int i) {
  return make_int4(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
}

template <int KMAX>
struct VecLds {
  static constexpr int LDV = KMAX + 1;
  static constexpr int RING = 8 * KMAX;              // rotation ring (entries)
  static constexpr int REFILL = RING / 2;
  static constexpr int A = 16 * LDV;                 // 16 reflectors
  static constexpr int B = 2 * (RING + 8);          // rotation ring + its mirrored head
  static constexpr int C = 16 * LDV;                 // 16 columns of V
  static constexpr int N = (A > B ? (A > C ? A : C) : (B > C ? B : C));
};

// offset (in doubles) of tile (I, J) register r of lane l in a trans block
template <int KMAX>
__device__ __forceinline__ int tile_off(int I, int J, int r, int l) {
  return ((I * (KMAX / 16) + J) * 4 + r) * 64 + l;
}

template <int KMAX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_vec(EigParams p) {
  constexpr int NB = KMAX / 8;
  constexpr int TT = KMAX / 16;
  constexpr int IPP = TT < 2 ? TT : 2;               // tile rows per pass
  using L = VecLds<KMAX>;
  constexpr int LDV = L::LDV, RING = L::RING, REFILL = L::REFILL;
  static_assert(REFILL >= KMAX + 16 && REFILL % 64 == 0, "a refill covers any chase");
  __shared__ __align__(16) double sbuf[L::N];
  // stage C's vectors share one array with stage A's Gram matrices (16 x 16),
  // which are dead by then: k_vec<64> then needs 11.4 KB of LDS, so eight of
  // its workgroups and one of k_ql_log's (66.5 KB) fit a CU together
  constexpr int NSC = 5 * KMAX > 256 ? 5 * KMAX : 256;
  __shared__ double ssc[NSC], stau[KMAX];
  double* const sf = ssc;
  double* const sg = ssc + KMAX;
  double* const sb = ssc + 2 * KMAX;
  double* const st = ssc + 3 * KMAX;
  double* const sw = ssc + 4 * KMAX;
  double* const sgram = ssc;
  const int lane = threadIdx.x;
  long long a, ai, ar;
  if (!eig_slot(p, blockIdx.x, a, ai, ar)) return;
  if (p.status[a] != 0) {                // redone by the list pass, or failed
    if (p.list || p.status[a] == kStatusNotRedone) nan_transform(p, a);
    return;
  }
  const int k = p.k, kp = p.kp;
  const bool in_r = lane < k;
  double* __restrict__ refl = p.refl + ai * (long long)kp * kp;
  const double* __restrict__ tg = p.tau + ai * kp;
  const int nref = k > 2 ? k - 2 : 0;

  // ---- A. Z <- Z H_j, reflectors staged 16 at a time ----
  // The next 16 reflectors are loaded into registers while the current 16
  // are applied, so their memory latency is hidden behind the work.
  if (lane < nref) stau[lane] = tg[lane];
  double z[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) z[c] = (c == lane) ? 1.0 : 0.0;
  double vnext[16];
#pragma unroll
  for (int t = 0; t < 16; ++t)
    vnext[t] = (t < nref && lane < KMAX) ? refl[(long long)t * kp + lane] : 0.0;
  // Reflectors apply in groups of four: the four dots from the same z, each
  // corrected by the earlier ones' scales times the pairwise products
  // v_s . v_t (the dot with z as the earlier reflectors of the group left
  // it), then one pass z -= sum_t alpha_t v_t.  A group is one dependent
  // step instead of four.  The products v_s . v_t of a staged block are its
  // Gram matrix V^T V, one f64 MFMA 16x16x4 per four columns.
  for (int j0 = 0; j0 < nref; j0 += 16) {
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; ++t)
      if (lane < KMAX) sbuf[t * LDV + lane] = vnext[t];
    __syncthreads();
    {
      const int i16 = lane & 15, q = lane >> 4;
      d4 g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s4 = 0; s4 < KMAX / 4; ++s4) {
        const double v = sbuf[i16 * LDV + 4 * s4 + q];      // V[4 s4 + q][i16]
        g = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, g, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) sgram[(q + 4 * r) * 16 + i16] = g[r];
    }
#pragma unroll
    for (int t = 0; t < 16; ++t)
      vnext[t] = (j0 + 16 + t < nref && lane < KMAX) ? refl[(long long)(j0 + 16 + t) * kp + lane]
                                                     : 0.0;
    __syncthreads();
    const int j1 = min(j0 + 16, nref);
    for (int j = j0; j < j1; j += 4) {
      // past nref the staged columns are zero (tau 0), so a short last group
      // runs with zero reflectors
      const int t0 = j - j0;
      double tau[4], al[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) tau[u] = j + u < j1 ? uniform(stau[j + u]) : 0.0;
      // the group's reflector entries are the same for every lane: they are
      // read as scalars (s_load from the reflector block, rows past the last
      // reflector from a zero row) and enter the FMAs as scalar operands, so
      // the group's two passes take no LDS bandwidth
      uniform_dbl_cptr qr[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        qr[u] = j + u < j1 ? (uniform_dbl_cptr)(refl + (long long)(j + u) * kp)
                           : (uniform_dbl_cptr)kZeroRow;
      double acc[4][2];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u][0] = acc[u][1] = 0.0;
#pragma unroll
      for (int B = 0; B < NB; ++B) {
        if (8 * B + 7 <= j) continue;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int c = 8 * B + t;
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[u][t & 1] = fma(z[c], qr[u][c], acc[u][t & 1]);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        double d = acc[u][0] + acc[u][1];
#pragma unroll
        for (int v = 0; v < u; ++v) d = fma(-al[v], uniform(sgram[(t0 + v) * 16 + t0 + u]), d);
        al[u] = d * tau[u];
      }
#pragma unroll
      for (int B = 0; B < NB; ++B) {
        if (8 * B + 7 <= j) continue;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int c = 8 * B + t;
          double zc = z[c];
#pragma unroll
          for (int u = 0; u < 4; ++u) zc = fma(-al[u], qr[u][c], zc);
          z[c] = zc;
        }
      }
    }
  }
  __syncthreads();

  // ---- B. logged QL rotations streamed through an LDS ring of REFILL-entry refills ----
  {
    double2* sring = reinterpret_cast<double2*>(sbuf);
    const double2* __restrict__ rot = p.rot + ar * (long long)p.rot_cap;
    // the chase headers are read as scalars (s_load), so waiting for one does
    // not wait for the vector loads of the next refill issued before it
    const uniform_int_cptr hdr =
        (uniform_int_cptr)(p.hdr + (long long)uniform((int)ar) * p.hdr_cap);
    const int nch = uniform(min(p.nchase[ar], p.hdr_cap));
    int nrot = 0;
    if (nch > 0) {
      const int4 hl = load_uniform_int4(hdr, nch - 1);
      nrot = uniform(hl.z + hl.w);
    }
    int committed = 0;
    int4 hn = nch > 0 ? load_uniform_int4(hdr, 0) : make_int4(0, 0, 0, 0);
    // the entries of the next refill, loaded one refill ahead so that the
    // log's memory latency is hidden behind the chases in between
    double2 vnext[REFILL / 64];
#pragma unroll
    for (int t = 0; t < REFILL / 64; ++t) {
      const int e = 64 * t + lane;
      vnext[t] = (e < nrot) ? rot[e] : make_double2(1.0, 0.0);
    }
    for (int ch = 0; ch < nch; ++ch) {
      int4 h = hn;
      h.x = uniform(h.x);
      h.y = uniform(h.y);
      h.z = uniform(h.z);
      h.w = uniform(h.w);
      if (ch + 1 < nch) hn = load_uniform_int4(hdr, ch + 1);
      const int lo = h.x, hi = h.y;
      if (h.z + h.w > committed) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < REFILL / 64; ++t) {
          const int slot = (committed + 64 * t + lane) & (RING - 1);
          sring[slot] = vnext[t];
          if (slot < 8) sring[RING + slot] = vnext[t];
        }
        committed += REFILL;
#pragma unroll
        for (int t = 0; t < REFILL / 64; ++t) {
          const int e = committed + 64 * t + lane;
          vnext[t] = (e < nrot) ? rot[e] : make_double2(1.0, 0.0);
        }
        __syncthreads();
      }
      const int bhi = (hi - 1) >> 3, blo = lo >> 3;
      const int base = h.z + ((hi + 7) & ~7) - 1;
      double carry = 0.0;
#pragma unroll
      for (int B = NB - 1; B >= 0; --B) {
        if (B > bhi || B < blo) continue;
        const int top = (8 * B + 7 < KMAX - 1) ? 8 * B + 7 : KMAX - 2;
        if (B == bhi) carry = z[top + 1];
        double2 cs[8];
        ring_block<RING>(sring, base, B, cs);
#pragma unroll
        for (int t = 7; t >= 0; --t) {
          const int i = 8 * B + t;
          if (i > KMAX - 2) continue;
          const double zi = z[i];
          // carry is the serial chain: one FMA per step
          z[i + 1] = fma(cs[t].x, carry, cs[t].y * zi);
          carry = fma(-cs[t].y, carry, cs[t].x * zi);
        }
        if (B == blo) z[8 * B] = carry;
      }
    }
  }

  // ---- C. eigenvalue functions (letkf_core.F90:107-113, guard for reduced modes) ----
  double lam = in_r ? p.lam[ar * kp + lane] : 0.0;
  if (p.guard) {
    const double mn = wave_min_full(in_r ? lam : 1e300);
    if (mn < 0.0) lam = lam - mn;
  }
  lam = lam * 1.0 + (double)(k - 1) / p.inflation;
  if (lane < KMAX) {
    sf[lane] = in_r ? sqrt((double)(k - 1) / lam) : 0.0;
    sg[lane] = in_r ? 1.0 / lam : 0.0;
    sb[lane] = in_r ? p.b[ai * kp + lane] : 0.0;
  }
  // T = V diag(f) V^T on f64 MFMA in IPP tile rows per pass; V is staged
  // 16 columns at a time ([column][row] in LDS).  The first pass also forms
  // st = diag(1/lambda) V^T b (letkf_core.F90:129-195 regrouped).
  const int i16 = lane & 15, q = lane >> 4;
  double* sVc = sbuf;
  double wbar = 0.0;
#pragma unroll
  for (int pass = 0; pass < TT / IPP; ++pass) {
    d4 acc[IPP][TT];
#pragma unroll
    for (int I = 0; I < IPP; ++I)
#pragma unroll
      for (int J = 0; J < TT; ++J) acc[I][J] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < TT; ++c) {
      __syncthreads();
      if (lane < KMAX) {
#pragma unroll
        for (int t = 0; t < 16; ++t) sVc[t * LDV + lane] = z[16 * c + t];
      }
      __syncthreads();
      if (pass == 0 && lane < 16) {
        double tb = 0.0;
        for (int r = 0; r < KMAX; ++r) tb = fma(sVc[lane * LDV + r], sb[r], tb);
        st[16 * c + lane] = tb * sg[16 * c + lane];
      }
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int t = 4 * s4 + q;                      // column within the chunk
        const double fc = sf[16 * c + t];
        double fr[TT];
#pragma unroll
        for (int J = 0; J < TT; ++J) fr[J] = sVc[t * LDV + 16 * J + i16];
#pragma unroll
        for (int I = 0; I < IPP; ++I)
#pragma unroll
          for (int J = 0; J < TT; ++J)
            acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[pass * IPP + I], fc * fr[J],
                                                             acc[I][J], 0, 0, 0);
      }
    }
    if (pass == 0) {
      __syncthreads();
#pragma unroll
      for (int c = 0; c < KMAX; ++c) wbar = fma(z[c], st[c], wbar);
      if (lane < KMAX) sw[lane] = in_r ? wbar : 0.0;
      __syncthreads();
    }
    // trans(f, e) = T(f, e) + overline_w(f)  (letkf_core.F90:221-225)
#pragma unroll
    for (int I = 0; I < IPP; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double wf = sw[16 * (pass * IPP + I) + q + 4 * r];
#pragma unroll
        for (int J = 0; J < TT; ++J) acc[I][J][r] += wf;
      }
    if (p.trans_out) {
      double* tr = p.trans_out + a * (long long)k * k;
#pragma unroll
      for (int I = 0; I < IPP; ++I)
#pragma unroll
        for (int J = 0; J < TT; ++J)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int f = 16 * (pass * IPP + I) + q + 4 * r, e = 16 * J + i16;
            if (f < k && e < k) tr[(long long)e * k + f] = acc[I][J][r];
          }
    }
    if (p.xb || p.xb32) {
      // the B operand of the update, over the (consumed) reflector block; for
      // the FP32 state in FP32 (the update rounds it to FP32 anyway), half
      // the bytes to write here and to read there
      if (p.state_f32) {
        float* const rf = reinterpret_cast<float*>(refl);
#pragma unroll
        for (int I = 0; I < IPP; ++I)
#pragma unroll
          for (int J = 0; J < TT; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              rf[tile_off<KMAX>(pass * IPP + I, J, r, lane)] = (float)acc[I][J][r];
      } else {
#pragma unroll
        for (int I = 0; I < IPP; ++I)
#pragma unroll
          for (int J = 0; J < TT; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) refl[tile_off<KMAX>(pass * IPP + I, J, r, lane)] = acc[I][J][r];
      }
    }
  }
  if (__builtin_amdgcn_ballot_w64(in_r && !__builtin_isfinite(wbar))) {
    __syncthreads();                               // the trans writes above are done
    if (lane == 0) p.status[a] = kStatusNonFinite;
    nan_transform(p, a);
    return;
  }
  if (p.w_out && in_r) p.w_out[a * (long long)k + lane] = wbar;
}

// State precision of the update: FP64 (xb/xa) or FP32 (xb32/xa32) members on
// f64 or f32 MFMA 16x16x4 (the two share their operand layout).
template <class TS> struct StateT;
template <> struct StateT<double> {
  using acc = d4;
  static __device__ __forceinline__ const double* xb(const EigParams& p) { return p.xb; }
  static __device__ __forceinline__ double* xa(const EigParams& p) { return p.xa; }
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct StateT<float> {
  using acc = f4;
  static __device__ __forceinline__ const float* xb(const EigParams& p) { return p.xb32; }
  static __device__ __forceinline__ float* xa(const EigParams& p) { return p.xa32; }
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// One state row o (k members at o*k) of a fused-epilogue column: the
// background mean ab, the analysis members xe (k of them, FP64) with mean aa;
// writes xa_out and the two spreads (module_da_parallel.F90:222-258).
__device__ __forceinline__ long long spread_index(const EigParams& p, int lo, int la, int v) {
  return ((long long)(lo - 1) * p.g.nlat + (la - 1)) * p.n + v;
}

// An analysis whose eigensolve failed on both passes (or found no room in the
// list pass): its state rows keep the background -- xa = xb, axa = axb, or
// with the fused epilogue xa_out = xb + axb and spread_a = spread_b.
template <class TS, int EPI>
__device__ void copy_background_t(const EigParams& p, long long a) {
  const AnRows R = an_rows(p, a);
  const int k = p.k;
  const TS* __restrict__ xb = StateT<TS>::xb(p);
  for (int lo = R.si; lo <= R.ei; ++lo)
    for (int la = R.sj; la <= R.ej; ++la) {
      const long long off = ((long long)(lo - 1 + p.halo) * R.lat_ext + (la - 1 + p.halo)) * p.n;
      for (int vv = 0; vv < p.n_var; ++vv) {
        int l0, l1;
        const int start = p.var_start[vv] - 1;
        if (p.var_lev[vv] == 1) {
          if (p.vertical && R.ek != p.nz) continue;
          l0 = l1 = start;
        } else if (p.vertical) {
          l0 = start + R.sk - 1;
          l1 = start + R.ek - 1;
        } else {
          l0 = start;
          l1 = start + p.var_lev[vv] - 1;
        }
        for (int nk = l0 + (int)threadIdx.x; nk <= l1; nk += blockDim.x) {
          const long long o = off + nk;
          const double ab = p.axb[o];
          double sb = 0.0;
          for (int m = 0; m < k; ++m) {
            const double x = (double)xb[o * k + m];
            if (EPI) p.xa_out[o * k + m] = x + ab;
            else StateT<TS>::xa(p)[o * k + m] = xb[o * k + m];
            sb += x * x;
          }
          if (EPI) {
            const double sp = sqrt(sb / (double)(float)k);
            p.spread_b[spread_index(p, lo, la, nk)] = sp;
            p.spread_a[spread_index(p, lo, la, nk)] = sp;
          } else {
            p.axa[o] = ab;
          }
        }
      }
    }
}

__device__ void copy_background(const EigParams& p, long long a) {
  if (p.epi) {
    if (p.state_f32) copy_background_t<float, 1>(p, a);
    else copy_background_t<double, 1>(p, a);
  } else {
    if (p.state_f32) copy_background_t<float, 0>(p, a);
    else copy_background_t<double, 0>(p, a);
  }
}

// k_update: xa(:, v) = trans^T xb(:, v) + axb(v), then re-centred, for the
// variables v of one analysis (letkf_core.F90 ensemble update; the re-centring
// is the mean/perturbation split of letkf_run).  One block of four wavefronts
// per analysis: trans (k x k, from the B-operand tiles k_vec leaves in the
// reflector block) is staged once in LDS and wavefront w takes the
// (column, 16-variable group) items w, w + 4, ...  The MFMA reduction index
// is permuted -- at step s, lane quarter q carries member S q + s (S = KMAX/4)
// -- so that a lane's A operands for a group are one contiguous run of S
// members of one variable (16-byte loads); B takes the same permutation from
// LDS.  Registers stay near 100, so four blocks share a CU and the loads of
// one wavefront overlap the MFMAs of the others.
//
// TS is the storage precision of the perturbations (FP32: f32 MFMA with FP32
// accumulation; the means and the re-centring stay FP64).  EPI = 1 fuses the
// caller's epilogue into the output pass: full members gamma1 (axb + xb) +
// gamma2 (axa + xa) (the background where |lat| > polar_lat) and the spreads,
// written in place of xa/axa.
template <int KMAX, class TS, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_update(EigParams p) {
  using ST = StateT<TS>;
  using acc_t = typename ST::acc;
  constexpr int TT = KMAX / 16;
  constexpr int S = KMAX / 4;
  constexpr int VEC = 16 / sizeof(TS);       // members per 16-byte load
  // S * LDT = 16 (mod 32) for FP64: the four lane quarters of a B read fall
  // on disjoint banks in each half-wavefront
  constexpr int LDT = KMAX == 64 ? KMAX + 1 : KMAX + 2;
  static_assert(KMAX == 32 || KMAX == 64, "k_update covers k <= 64");
  __shared__ TS sT[KMAX * LDT];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  long long a, ai, ar;
  if (!eig_slot(p, blockIdx.x, a, ai, ar)) return;
  if (p.status[a] != 0) {                // redone by the list pass, or failed
    if (p.list || keeps_background(p.status[a])) copy_background_t<TS, EPI>(p, a);
    return;
  }
  const int k = p.k;
  const int i16 = lane & 15, q = lane >> 4;
  // the row of a 16x16 MFMA result that register r of lane quarter q holds:
  // q + 4 r for f64 16x16x4, 4 q + r for f32 16x16x4
  constexpr bool F32 = sizeof(TS) == 4;
  auto out_row = [q](int r) { return F32 ? 4 * q + r : q + 4 * r; };
  const TS* __restrict__ xb = ST::xb(p);
  // trans tiles from k_vec: FP32 for the FP32 state, else FP64
  const double* __restrict__ tb = p.refl + ai * (long long)p.kp * p.kp;
  const TS* __restrict__ tbs = reinterpret_cast<const TS*>(tb);
  for (int idx = tid; idx < TT * TT * 4 * 64; idx += 256) {
    const int l = idx & 63, c = idx >> 6, r = c & 3, IJ = c >> 2;
    const int row = 16 * (IJ / TT) + 4 * r + (l >> 4), col = 16 * (IJ % TT) + (l & 15);
    sT[row * LDT + col] = (row < k && col < k) ? tbs[idx] : (TS)0;
  }

  const AnRows R = an_rows(p, a);
  const int sj = R.sj, ej = R.ej, si = R.si, ei = R.ei, sk = R.sk, ek = R.ek;
  const long long lat_ext = R.lat_ext;
  // the analysis's state rows: all n of the column, or (vertical loop) the
  // levels sk..ek of each 3-D variable and the 2-D variables on the top chunk;
  // sidx maps the v-th of them to its offset in the column
  auto run_len = [&](int vv) {
    return p.var_lev[vv] == 1 ? (ek == p.nz ? 1 : 0) : ek - sk + 1;
  };
  int nvec = p.n;
  if (p.vertical) {
    nvec = 0;
    for (int vv = 0; vv < p.n_var; ++vv) nvec += run_len(vv);
  }
  auto sidx = [&](int v) {
    if (v >= nvec) return -1;
    if (!p.vertical) return v;
    int rem = v;
    for (int vv = 0; vv < p.n_var; ++vv) {
      const int len = run_len(vv);
      if (rem < len) return p.var_start[vv] - 1 + (p.var_lev[vv] == 1 ? 0 : sk - 1) + rem;
      rem -= len;
    }
    return -1;
  };
  const int ngrp = (nvec + 15) / 16;
  const int ncl = ej - sj + 1;
  const int n_items = (ei - si + 1) * ncl * ngrp;
  const bool vec_ld = (k == KMAX) && ((reinterpret_cast<uintptr_t>(xb) & 15) == 0);
  auto item_off = [&](int it) {
    const int cl = it / ngrp;
    const int lo = si + cl / ncl, la = sj + cl % ncl;
    return ((long long)(lo - 1 + p.halo) * lat_ext + (la - 1 + p.halo)) * p.n;
  };
  // A operands of item it: variable 16 g + i16, members S q .. S q + S - 1
  auto load_a = [&](int it, TS (&af)[S]) {
    const long long off = item_off(it);
    const int s_a = sidx(16 * (it % ngrp) + i16);
    if (s_a >= 0 && vec_ld) {
      const uint4* xr = reinterpret_cast<const uint4*>(xb + (off + s_a) * k + S * q);
#pragma unroll
      for (int u = 0; u < S / VEC; ++u) {
        const uint4 t4 = xr[u];
        const TS* tv = reinterpret_cast<const TS*>(&t4);
#pragma unroll
        for (int x = 0; x < VEC; ++x) af[VEC * u + x] = tv[x];
      }
    } else {
      const TS* xr = xb + (off + (s_a >= 0 ? s_a : 0)) * k;
#pragma unroll
      for (int u = 0; u < S; ++u) {
        const int m = S * q + u;
        af[u] = (s_a >= 0 && m < k) ? xr[m] : (TS)0;
      }
    }
  };
  // the first item's A operands are read before the staging barrier, each
  // later item's as soon as the previous item's MFMAs have consumed theirs
  TS af[S];
  if (w < n_items) load_a(w, af);
  __syncthreads();
  for (int it = w; it < n_items; it += 4) {
    const long long off = item_off(it);
    const int v0 = 16 * (it % ngrp);
    double ab4[4];
    int st4[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      st4[r] = sidx(v0 + out_row(r));
      ab4[r] = (st4[r] >= 0) ? p.axb[off + st4[r]] : 0.0;
    }
    acc_t out[TT];
#pragma unroll
    for (int J = 0; J < TT; ++J) out[J] = acc_t{0, 0, 0, 0};
    // B of step s + 1 is read while the MFMAs of step s run; the scheduling
    // barriers keep the compiler from hoisting every B read (and its
    // registers) to the top of the loop
    const TS* bq = sT + S * q * LDT + i16;
    TS bn[TT];
#pragma unroll
    for (int J = 0; J < TT; ++J) bn[J] = bq[16 * J];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      TS b[TT];
#pragma unroll
      for (int J = 0; J < TT; ++J) b[J] = bn[J];
      if (s + 1 < S) {
#pragma unroll
        for (int J = 0; J < TT; ++J) bn[J] = bq[(s + 1) * LDT + 16 * J];
      }
#pragma unroll
      for (int J = 0; J < TT; ++J) out[J] = ST::mma(af[s], b[J], out[J]);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (it + 4 < n_items) load_a(it + 4, af);
    // lane (i16, q) holds xa of variable v0 + out_row(r), member 16 J + i16
    int lo_c = 0, la_c = 0;
    bool polar = false;
    if (EPI) {
      const int cl = it / ngrp;
      lo_c = si + cl / ncl;
      la_c = sj + cl % ncl;
      polar = fabs((double)p.g.glat[la_c - 1]) > p.polar_lat;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double sum = 0.0;
      double xe[TT];
#pragma unroll
      for (int J = 0; J < TT; ++J) {
        const int e = 16 * J + i16;
        xe[J] = (e < k) ? (double)out[J][r] + ab4[r] : 0.0;
        sum += xe[J];
      }
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) sum += __shfl_xor(sum, o, 64);
      const double mean = sum / k;
      if (!EPI) {
        if (st4[r] >= 0) {
          const long long o = off + st4[r];
          TS* __restrict__ xa = ST::xa(p);
#pragma unroll
          for (int J = 0; J < TT; ++J) {
            const int e = 16 * J + i16;
            if (e < k) xa[o * k + e] = (TS)(xe[J] - mean);
          }
          if (i16 == 0) p.axa[o] = mean;
        }
      } else {
        // the caller's epilogue (module_da_parallel.F90:219-261)
        const long long o = off + (st4[r] >= 0 ? st4[r] : 0);
        double sb = 0.0, sa = 0.0;
        double xo[TT];
#pragma unroll
        for (int J = 0; J < TT; ++J) {
          const int e = 16 * J + i16;
          const bool in = e < k && st4[r] >= 0;
          const double xbe = in ? (double)xb[o * k + e] : 0.0;
          const double xae = xe[J] - mean;
          double v;
          if (polar) {
            v = xbe + ab4[r];
            sa += xbe * xbe;
          } else {
            v = p.gamma1 * (ab4[r] + xbe) + p.gamma2 * (mean + xae);
            if (e < k) sa += (v - mean) * (v - mean);
          }
          xo[J] = v;
          sb += xbe * xbe;
        }
#pragma unroll
        for (int o2 = 1; o2 < 16; o2 <<= 1) {
          sb += __shfl_xor(sb, o2, 64);
          sa += __shfl_xor(sa, o2, 64);
        }
        if (st4[r] >= 0) {
#pragma unroll
          for (int J = 0; J < TT; ++J) {
            const int e = 16 * J + i16;
            if (e < k) p.xa_out[o * k + e] = xo[J];
          }
          if (i16 == 0) {
            const long long so = spread_index(p, lo_c, la_c, st4[r]);
            p.spread_b[so] = sqrt(sb / (double)(float)k);
            p.spread_a[so] = sqrt(sa / (double)(float)k);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_vec128 + k_update128: k = 65..128 with the rows of Z in registers.
// 256 threads per matrix: thread t keeps half h = t / 128 (columns
// 64h .. 64h+63) of row r = t % 128.  A chase that crosses the middle (lo < 64
// <= hi) runs on the upper half first; its carry goes to the lower half
// through LDS, and the lower half's step at position 63 returns the new
// column-64 entry the same way.  The lower part runs one chase late, beside
// the next chase's upper part, so both halves work through runs of crossing
// chases.
// ---------------------------------------------------------------------------

// Householder tridiagonalisation for k = 65..128 with half rows of A in
// registers (thread t: row t % 128, the columns of parity h = t / 128, one
// half per wavefront): the halves interleave, so both keep the same share of
// the shrinking trailing matrix, and blocks of columns <= j (where v and w
// are zero) are skipped.  A column already in tridiagonal form runs the same
// step with v = 0 and tau = 0 (an exact identity update), so the loop has one
// path and A stays in the same registers.  Three wavefronts per SIMD.  Same
// outputs as k_tridiag.  Three barriers per step:
//  - the pivot column j is in LDS (double-buffered by the parity of j) with
//    rows <= j zeroed; each wavefront writes v_{j+1} = x0 - alpha over its
//    entry j+1 before reading it (every wavefront writes the same value), so
//    v is read from the column directly; x0 and the sum of squares of the
//    column below it come from the previous step;
//  - the partial products of the two halves and the per-wavefront partials of
//    v.p share one barrier;
//  - w, then the update, the extraction of column j+1 (stores under
//    wavefront-uniform branches, read back for its squares) share the step's
//    last barrier.
// Partial sums are combined in a fixed order, so every thread holds the same
// values.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_tridiag128(EigParams p) {
  constexpr int KMAX = 128;
  __shared__ double sx[2][KMAX];          // pivot column, rows <= j zeroed
  __shared__ double sx0[2];               // A(j+1, j) of the pivot column
  __shared__ double sw[KMAX], spr[2][KMAX];
  __shared__ double se_s[KMAX + 1], st_s[KMAX + 1], sd[KMAX];
  __shared__ double rsq[2][4], rvp[4];    // per-wavefront partials
  const int tid = threadIdx.x;
  const int r = tid & 127, h = uniform(tid >> 7);   // the half is one per wavefront
  const int wv = tid >> 6, lane = tid & 63;
  const long long a = p.a0 + blockIdx.x;
  if (a >= p.a0 + p.n_an) return;
  const long long al = a - p.a0;
  const int k = p.k, kp = p.kp;
  const double* __restrict__ Ag = p.A + al * (long long)kp * kp;
  double* __restrict__ refl = p.refl + al * (long long)kp * kp;
  const bool in_r = r < k;
  double arow[64];
  // eight columns in flight at a time: the loads need one address each
#pragma unroll
  for (int B = 0; B < 8; ++B) {
#pragma unroll
    for (int c = 8 * B; c < 8 * B + 8; ++c) {
      // unconditional in-bounds load (r < kp), then a select: no branch
      // per column
      const int col = 2 * c + h;
      const double av = Ag[(long long)col * kp + r];
      arow[c] = (in_r && col < k) ? av : 0.0;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  // column 0 (held by half 0) and the squares of its rows > 1
  if (h == 0) {
    sx[0][r] = (r > 0) ? arow[0] : 0.0;
    if (r == 1) sx0[0] = arow[0];
  }
  {
    const double sq = wave_sum_full((h == 0 && r > 1) ? arow[0] * arow[0] : 0.0);
    if (lane == 0) rsq[0][wv] = sq;
  }
  __syncthreads();
  // column j+1 -> sx[nxt] (rows <= j+1 zeroed) and sx0[nxt]; partial squares
  // of its rows > j+2 -> rsq[nxt]
  auto pass_column = [&](int j, int nxt) {
    const int hj1 = (j + 1) & 1;
    if (h == hj1) {
      const double xc = pick_column(arow, (j + 1) >> 1);
      sx[nxt][r] = (r > j + 1) ? xc : 0.0;
      if (r == j + 2) sx0[nxt] = xc;
    }
    const double xn = (h == hj1 && r > j + 2) ? sx[nxt][r] : 0.0;
    const double sq = wave_sum_full(xn * xn);
    if (lane == 0) rsq[nxt][wv] = sq;
  };
  for (int j = 0; j + 2 < k; ++j) {
    const int cur = j & 1, nxt = cur ^ 1;
    double* col = sx[cur];
    const bool act = in_r && r > j;
    const double x0 = uniform(sx0[cur]);
    const double rest = (rsq[cur][0] + rsq[cur][1]) + (rsq[cur][2] + rsq[cur][3]);
    // a column already in tridiagonal form (rest = 0) takes the same path
    // with v = 0 and tau = 0: the update is then exactly the identity
    const bool none = rest == 0.0;
    const double nrm = sqrt(rest + x0 * x0);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;
    const double v0 = none ? 0.0 : x0 - alpha;
    const double tau = none ? 0.0 : 2.0 / (rest + v0 * v0);
    if (lane == 0) col[j + 1] = v0;
    const double vr = act ? col[r] : 0.0;
    // partial p over this half (v_c = 0 for c <= j)
    double p0 = 0.0, p1 = 0.0;
#pragma unroll
    for (int B = 0; B < 8; ++B) {
      if (16 * B + 14 + h <= j) continue;      // v is zero over these columns
#pragma unroll
      for (int c = 8 * B; c < 8 * B + 8; c += 2) {
        p0 = fma(arow[c], col[2 * c + h], p0);
        p1 = fma(arow[c + 1], col[2 * c + 2 + h], p1);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    const double ph = p0 + p1;
    if (in_r) spr[h][r] = ph;
    {
      const double vp = wave_sum_full(vr * ph);
      if (lane == 0) rvp[wv] = vp;
    }
    __syncthreads();
    const double pr = act ? (spr[0][r] + spr[1][r]) * tau : 0.0;
    const double K = 0.5 * tau * tau * ((rvp[0] + rvp[1]) + (rvp[2] + rvp[3]));
    const double wr = act ? pr - K * vr : 0.0;
    if (h == 0) sw[r] = wr;
    __syncthreads();
    // A -= v w^T + w v^T over this half (vr = wr = 0 in rows <= j)
#pragma unroll
    for (int B = 0; B < 8; ++B) {
      if (16 * B + 14 + h <= j) continue;      // v and w are zero over these columns
#pragma unroll
      for (int c = 8 * B; c < 8 * B + 8; ++c) {
        const int cg = 2 * c + h;
        arow[c] = fma(-wr, col[cg], fma(-vr, sw[cg], arow[c]));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    pass_column(j, nxt);
    if (h == 0) refl[(long long)j * kp + r] = vr;
    if (tid == 0) { se_s[j] = none ? x0 : alpha; st_s[j] = tau; }
    __syncthreads();
  }
  // diagonal (thread of the half holding column r) and e[k-2] = A[k-1][k-2]
#pragma unroll
  for (int c = 0; c < 64; ++c)
    if (2 * c + h == r) sd[r] = arow[c];
  if (r == k - 1) {
#pragma unroll
    for (int c = 0; c < 64; ++c)
      if (2 * c + h == k - 2) se_s[KMAX] = arow[c];
  }
  __syncthreads();
  if (h != 0) return;
  double* __restrict__ dg = p.d + al * kp;
  double* __restrict__ eg = p.e + al * kp;
  double* __restrict__ tg = p.tau + al * kp;
  dg[r] = in_r ? sd[r] : 0.0;
  double ev = 0.0, tv = 0.0;
  if (r + 2 < k) { ev = se_s[r]; tv = st_s[r]; }
  else if (r == k - 2) ev = se_s[KMAX];
  eg[r] = ev;
  tg[r] = tv;
}

// Applies the positions il in [lo_l, hi_l) of a chase to a half row
// z[0..63] (rotation of local position il at ring entry base_l - il, identity
// outside); z[64] is `ztop` (in/out).  The carry enters at the top block as
// z[top+1] and leaves as the return value; with `final_write`, z[lo_l] takes
// it (the chase ends in this half).
template <int RING>
__device__ __forceinline__ double chase_half(double (&z)[64], double& ztop, int lo_l, int hi_l,
                                             const double2* sring, int base_l, bool final_write,
                                             bool top_done) {
  const int bhi = (hi_l - 1) >> 3, blo = lo_l >> 3;
  double carry = 0.0;
#pragma unroll
  for (int B = 7; B >= 0; --B) {
    if (B > bhi || B < blo) continue;
    if (B == bhi) carry = (B == 7) ? ztop : z[8 * B + 8];
    double2 cs[8];
    ring_block<RING>(sring, base_l, B, cs);
    if (B == 7 && top_done) cs[7] = make_double2(1.0, 0.0);
#pragma unroll
    for (int t = 7; t >= 0; --t) {
      const int i = 8 * B + t;
      const double zi = z[i];
      const double up = fma(cs[t].x, carry, cs[t].y * zi);
      if (i == 63) ztop = up;
      else z[i + 1] = up;
      carry = fma(-cs[t].y, carry, cs[t].x * zi);
    }
    if (B == blo && final_write) z[8 * B] = carry;
  }
  return carry;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_vec128(EigParams p) {
  constexpr int KMAX = 128, LDV = KMAX + 1, NT = 256;
  constexpr int RING = 1024, REFILL = 512;
  __shared__ __align__(16) double sbuf[16 * LDV > 2 * (RING + 8) ? 16 * LDV : 2 * (RING + 8)];
  __shared__ double sf[KMAX], sg[KMAX], sb[KMAX], st[KMAX], sw[KMAX], stau[KMAX];
  __shared__ double sx[2][2][KMAX];      // partial dots [parity][half][row]
  __shared__ double sc[KMAX], sz64[KMAX];
  __shared__ double red[8];
  const int tid = threadIdx.x;
  const int r = tid & 127, h = uniform(tid >> 7);   // the half is one per wavefront
  const int lane = tid & 63, wv = uniform(tid >> 6);
  long long a, ai, ar;
  if (!eig_slot(p, blockIdx.x, a, ai, ar)) return;
  if (p.status[a] != 0) {                // redone by the list pass, or failed
    if (p.list || p.status[a] == kStatusNotRedone) nan_transform(p, a);
    return;
  }
  const int k = p.k, kp = p.kp;
  const bool in_r = r < k;
  double* __restrict__ refl = p.refl + ai * (long long)kp * kp;
  const double* __restrict__ tg = p.tau + ai * kp;
  const int nref = k > 2 ? k - 2 : 0;

  // ---- A. Z <- Z H_j: half dots exchanged through LDS ----
  if (tid < nref) stau[tid] = tg[tid];
  double z[64];
#pragma unroll
  for (int c = 0; c < 64; ++c) z[c] = (64 * h + c == r) ? 1.0 : 0.0;
  for (int j0 = 0; j0 < nref; j0 += 16) {
    __syncthreads();
    for (int e = tid; e < 16 * KMAX; e += NT) {
      const int t = e / KMAX, c = e % KMAX;
      sbuf[t * LDV + c] = (j0 + t < nref) ? refl[(long long)(j0 + t) * kp + c] : 0.0;
    }
    __syncthreads();
    const int j1 = min(j0 + 16, nref);
    for (int j = j0; j < j1; ++j) {
      const double tau = uniform(stau[j]);
      const double* q = sbuf + (j - j0) * LDV + 64 * h;
      // the reflector is zero over columns <= j: blocks of 8 there are skipped
      double d0 = 0.0, d1 = 0.0;
#pragma unroll
      for (int B = 0; B < 8; ++B) {
        if (64 * h + 8 * B + 7 <= j) continue;
#pragma unroll
        for (int c = 8 * B; c < 8 * B + 8; c += 2) {
          d0 = fma(z[c], q[c], d0);
          d1 = fma(z[c + 1], q[c + 1], d1);
        }
      }
      sx[j & 1][h][r] = d0 + d1;
      __syncthreads();
      const double dot = (sx[j & 1][0][r] + sx[j & 1][1][r]) * tau;
#pragma unroll
      for (int B = 0; B < 8; ++B) {
        if (64 * h + 8 * B + 7 <= j) continue;
#pragma unroll
        for (int c = 8 * B; c < 8 * B + 8; ++c) z[c] = fma(-dot, q[c], z[c]);
      }
    }
  }
  __syncthreads();

  // ---- B. logged QL rotations ----
  {
    double2* sring = reinterpret_cast<double2*>(sbuf);
    const double2* __restrict__ rot = p.rot + ar * (long long)p.rot_cap;
    const int4* __restrict__ hdr = p.hdr + ar * (long long)p.hdr_cap;
    const int nch = uniform(min(p.nchase[ar], p.hdr_cap));
    int nrot = 0;
    if (nch > 0) {
      const int4 hl = hdr[nch - 1];
      nrot = uniform(hl.z + hl.w);
    }
    int committed = 0;
    double ztop = 0.0;
    // A crossing chase's lower part runs one chase late, beside the next
    // chase's upper part: its step at position 63 (which needs the upper
    // part's carry and gives column 64 back) runs alone between two barriers,
    // the rest of it together with the next chase's piece in the other half.
    bool pend = false;
    int plo = 0, pbase = 0;
    for (int ch = 0; ch <= nch; ++ch) {
      const bool have = ch < nch;       // ch == nch: the last pending lower part
      int lo = 0, hi = 0, base = 0;
      bool cross = false;
      if (have) {
        const int4 hh = hdr[ch];
        lo = uniform(hh.x);
        hi = uniform(hh.y);
        const int off = uniform(hh.z), len = uniform(hh.w);
        // refills are the only block-wide barriers besides the crossing
        // chases': the slots a refill overwrites may still be read by a chase
        // that started below the previous refill's entries, hence the leading
        // barrier (the pending chase is the previous one: a refill, needed
        // only when this chase ends past the ring's entries, never reaches it)
        if (off + len > committed) {
          __syncthreads();
          for (int e = tid; e < REFILL; e += NT) {
            const int g = committed + e;
            const double2 v = (g < nrot) ? rot[g] : make_double2(1.0, 0.0);
            const int slot = g & (RING - 1);
            sring[slot] = v;
            if (slot < 8) sring[RING + slot] = v;
          }
          committed += REFILL;
          __syncthreads();
        }
        base = off + ((hi + 7) & ~7) - 1;
        cross = lo < 64 && hi >= 64;
      }
      if (pend) {
        // position 63 of the pending chase: carry from the upper part (sc),
        // new column 64 to the upper half (sz64); z[63] keeps the carry on
        if (h == 0) {
          const double cin = sc[r];
          const double2 cs = sring[(pbase - 63) & (RING - 1)];
          const double z63 = z[63];
          sz64[r] = fma(cs.x, cin, cs.y * z63);
          z[63] = fma(-cs.y, cin, cs.x * z63);
        }
        __syncthreads();
        if (h == 1) z[0] = sz64[r];
      }
      // piece 0: the rest of the pending lower part (its position 63 is done);
      // piece 1: this chase's piece in this half (a crossing chase: the upper
      // part, whose carry goes to sc)
      for (int t = 0; t < 2; ++t) {
        bool act = false, fw = true;
        int lo_l = 0, hi_l = 0, bl = 0;
        if (t == 0) {
          act = pend && h == 0;
          lo_l = plo;
          hi_l = 64;
          bl = pbase;
        } else if (cross) {
          act = h == 1 && hi > 64;
          hi_l = hi - 64;
          bl = base - 64;
          fw = false;
        } else if (have) {
          act = (hi <= 63 && h == 0) || (lo >= 64 && h == 1);
          lo_l = lo - 64 * h;
          hi_l = hi - 64 * h;
          bl = base - 64 * h;
        }
        if (act) {
          ztop = 0.0;
          const double cy = chase_half<RING>(z, ztop, lo_l, hi_l, sring, bl, fw, t == 0);
          if (t == 1 && cross) sc[r] = cy;
        }
      }
      // hi == 64: no position in the upper half, the carry is column 64
      if (cross && h == 1 && hi == 64) sc[r] = z[0];
      pend = cross;
      if (pend) {
        plo = lo;
        pbase = base;
        __syncthreads();
      }
    }
  }
  __syncthreads();

  // ---- C. eigenvalue functions, T = V diag(f) V^T on MFMA, overline_w ----
  if (tid < KMAX) {
    double lam = tid < k ? p.lam[ar * kp + tid] : 0.0;
    sf[tid] = lam;
  }
  __syncthreads();
  if (p.guard) {
    double mn = 1e300;
    for (int i = 0; i < k; ++i) mn = fmin(mn, sf[i]);
    __syncthreads();
    if (tid < k && mn < 0.0) sf[tid] -= mn;
    __syncthreads();
  }
  if (tid < KMAX) {
    const bool ok = tid < k;
    const double lam = ok ? sf[tid] * 1.0 + (double)(k - 1) / p.inflation : 1.0;
    sg[tid] = ok ? 1.0 / lam : 0.0;
    sb[tid] = ok ? p.b[ai * kp + tid] : 0.0;
    sf[tid] = ok ? sqrt((double)(k - 1) / lam) : 0.0;
  }
  // st = diag(1/lambda) V^T b: column sums over rows, V staged 16 columns at
  // a time ([column][row]); T tiles: wave w computes J in {2w, 2w+1}, in two
  // passes over I (4 tile rows each)
  const int i16 = lane & 15, q = lane >> 4;
  double* sVc = sbuf;
  double wpart = 0.0;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    d4 acc[4][2];
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
      for (int J = 0; J < 2; ++J) acc[I][J] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {                // chunk c: columns 16c .. 16c+15
      __syncthreads();
      if (h == (c >> 2)) {
#pragma unroll
        for (int t = 0; t < 16; ++t) sVc[t * LDV + r] = z[16 * (c & 3) + t];
      }
      __syncthreads();
      if (pass == 0 && tid < 16) {
        double tb = 0.0;
        for (int rr = 0; rr < KMAX; ++rr) tb = fma(sVc[tid * LDV + rr], sb[rr], tb);
        st[16 * c + tid] = tb * sg[16 * c + tid];
      }
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int t = 4 * s4 + q;
        const double fc = sf[16 * c + t];
        double fa[4], fb[2];
#pragma unroll
        for (int I = 0; I < 4; ++I) fa[I] = sVc[t * LDV + 16 * (4 * pass + I) + i16];
#pragma unroll
        for (int J = 0; J < 2; ++J) fb[J] = fc * sVc[t * LDV + 16 * (2 * wv + J) + i16];
#pragma unroll
        for (int I = 0; I < 4; ++I)
#pragma unroll
          for (int J = 0; J < 2; ++J)
            acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[I], fb[J], acc[I][J], 0, 0, 0);
      }
    }
    if (pass == 0) {
      __syncthreads();
      // overline_w(r) = sum_c V[r][c] st_c, halves summed through LDS
#pragma unroll
      for (int c = 0; c < 64; ++c) wpart = fma(z[c], st[64 * h + c], wpart);
      sx[0][h][r] = wpart;
      __syncthreads();
      if (tid < KMAX) sw[tid] = tid < k ? sx[0][0][tid] + sx[0][1][tid] : 0.0;
      __syncthreads();
    }
    // trans(f, e) = T(f, e) + overline_w(f)
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const double wf = sw[16 * (4 * pass + I) + q + 4 * rr];
#pragma unroll
        for (int J = 0; J < 2; ++J) acc[I][J][rr] += wf;
      }
    if (p.trans_out) {
      double* tr = p.trans_out + a * (long long)k * k;
#pragma unroll
      for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = 0; J < 2; ++J)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const int f = 16 * (4 * pass + I) + q + 4 * rr, e = 16 * (2 * wv + J) + i16;
            if (f < k && e < k) tr[(long long)e * k + f] = acc[I][J][rr];
          }
    }
    if (p.xb || p.xb32) {
      __syncthreads();                             // every read of the reflector block is done
#pragma unroll
      for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = 0; J < 2; ++J)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr)
            refl[tile_off<KMAX>(4 * pass + I, 2 * wv + J, rr, lane)] = acc[I][J][rr];
    }
  }
  // every thread scans overline_w in LDS, so the decision is uniform
  bool bad = false;
  for (int f = 0; f < k; ++f) bad = bad || !__builtin_isfinite(sw[f]);
  if (bad) {
    __syncthreads();                               // the trans writes above are done
    if (threadIdx.x == 0) p.status[a] = kStatusNonFinite;
    nan_transform(p, a);
    return;
  }
  if (p.w_out && h == 0 && in_r) p.w_out[a * (long long)k + r] = sw[r];
}

// Xa = Xb trans for k = 65..128: wave w produces members 32w .. 32w+31 from the
// trans tiles J in {2w, 2w+1}; 16 state vectors per MFMA group.  TS and EPI
// as k_update.
template <class TS, int EPI>
__global__ __launch_bounds__(256) void k_update128(EigParams p) {
  using ST = StateT<TS>;
  using acc_t = typename ST::acc;
  constexpr int KMAX = 128, LDV = KMAX + 1, NT = 256;
  __shared__ TS sV[16 * LDV];
  __shared__ double spart[3][4][16];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  long long a, ai, ar;
  if (!eig_slot(p, blockIdx.x, a, ai, ar)) return;
  if (p.status[a] != 0) {                // redone by the list pass, or failed
    if (p.list || keeps_background(p.status[a])) copy_background_t<TS, EPI>(p, a);
    return;
  }
  const int k = p.k, kp = p.kp;
  const int i16 = lane & 15, q = lane >> 4;
  // the row of a 16x16 MFMA result that register rr of lane quarter q holds:
  // q + 4 rr for f64 16x16x4, 4 q + rr for f32 16x16x4
  constexpr bool F32 = sizeof(TS) == 4;
  auto out_row = [q](int rr) { return F32 ? 4 * q + rr : q + 4 * rr; };
  const TS* __restrict__ xb = ST::xb(p);
  const double* __restrict__ tb = p.refl + ai * (long long)kp * kp;
  TS bop[8][2][4];
#pragma unroll
  for (int I = 0; I < 8; ++I)
#pragma unroll
    for (int J = 0; J < 2; ++J)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) bop[I][J][rr] = (TS)tb[tile_off<KMAX>(I, 2 * wv + J, rr, lane)];

  const AnRows R = an_rows(p, a);
  const int sj = R.sj, ej = R.ej, si = R.si, ei = R.ei, sk = R.sk, ek = R.ek;
  const long long lat_ext = R.lat_ext;
  int rs[kMaxVars], rl[kMaxVars];
  int nrun = 0, nvec = 0;
  if (!p.vertical) {
    rs[0] = 0; rl[0] = p.n; nrun = 1;
  } else {
    for (int vv = 0; vv < p.n_var; ++vv) {
      const int start = p.var_start[vv] - 1;
      if (p.var_lev[vv] == 1) {
        if (ek == p.nz) { rs[nrun] = start; rl[nrun] = 1; ++nrun; }
      } else {
        rs[nrun] = start + sk - 1; rl[nrun] = ek - sk + 1; ++nrun;
      }
    }
  }
  for (int t = 0; t < nrun; ++t) nvec += rl[t];
  auto sidx = [&](int v) {
    int rem = v;
    for (int t = 0; t < nrun; ++t) {
      if (rem < rl[t]) return rs[t] + rem;
      rem -= rl[t];
    }
    return -1;
  };
  for (int lo = si; lo <= ei; ++lo) {
    for (int la = sj; la <= ej; ++la) {
      const long long off = ((long long)(lo - 1 + p.halo) * lat_ext + (la - 1 + p.halo)) * p.n;
      const bool polar = EPI && fabs((double)p.g.glat[la - 1]) > p.polar_lat;
      for (int v0 = 0; v0 < nvec; v0 += 16) {
        __syncthreads();
        for (int e = tid; e < 16 * KMAX; e += NT) {
          const int u = e / KMAX, m = e % KMAX;
          const int s_i = (v0 + u < nvec) ? sidx(v0 + u) : -1;
          sV[u * LDV + m] = (s_i >= 0 && m < k) ? xb[(off + s_i) * k + m] : (TS)0;
        }
        __syncthreads();
        acc_t out[2];
#pragma unroll
        for (int J = 0; J < 2; ++J) out[J] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int sK = 0; sK < KMAX / 4; ++sK) {
          const TS xa_frag = sV[i16 * LDV + 4 * sK + q];
#pragma unroll
          for (int J = 0; J < 2; ++J) out[J] = ST::mma(xa_frag, bop[sK >> 2][J][sK & 3], out[J]);
        }
        // lane holds Xa[vector out_row(rr)][member 32w + 16J + i16]
        double ab4[4], xe[4][2];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int v = v0 + out_row(rr);
          const int s_i = (v < nvec) ? sidx(v) : -1;
          ab4[rr] = (s_i >= 0) ? p.axb[off + s_i] : 0.0;
          double sum = 0.0;
#pragma unroll
          for (int J = 0; J < 2; ++J) {
            const int e = 32 * wv + 16 * J + i16;
            xe[rr][J] = (e < k) ? (double)out[J][rr] + ab4[rr] : 0.0;
            sum += xe[rr][J];
          }
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) sum += __shfl_xor(sum, o, 64);
          if (i16 == 0) spart[0][wv][out_row(rr)] = sum;
        }
        __syncthreads();
        if (!EPI) {
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const int v = v0 + out_row(rr);
            const int s_i = (v < nvec) ? sidx(v) : -1;
            if (s_i < 0) continue;
            const int vi = out_row(rr);
            const double mean =
                ((spart[0][0][vi] + spart[0][1][vi]) + (spart[0][2][vi] + spart[0][3][vi])) / k;
            const long long o = off + s_i;
            TS* __restrict__ xa = ST::xa(p);
#pragma unroll
            for (int J = 0; J < 2; ++J) {
              const int e = 32 * wv + 16 * J + i16;
              if (e < k) xa[o * k + e] = (TS)(xe[rr][J] - mean);
            }
            if (wv == 0 && i16 == 0) p.axa[o] = mean;
          }
        } else {
          // the caller's epilogue (module_da_parallel.F90:219-261); the
          // spreads' member sums go through LDS like the mean
          int sidx4[4];
          double mean4[4];
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const int v = v0 + out_row(rr);
            sidx4[rr] = (v < nvec) ? sidx(v) : -1;
            const int vi = out_row(rr);
            mean4[rr] =
                ((spart[0][0][vi] + spart[0][1][vi]) + (spart[0][2][vi] + spart[0][3][vi])) / k;
            const long long o = off + (sidx4[rr] >= 0 ? sidx4[rr] : 0);
            double sb = 0.0, sa = 0.0;
#pragma unroll
            for (int J = 0; J < 2; ++J) {
              const int e = 32 * wv + 16 * J + i16;
              const bool in = e < k && sidx4[rr] >= 0;
              const double xbe = in ? (double)xb[o * k + e] : 0.0;
              double vv;
              if (polar) {
                vv = xbe + ab4[rr];
                sa += xbe * xbe;
              } else {
                vv = p.gamma1 * (ab4[rr] + xbe) + p.gamma2 * (mean4[rr] + (xe[rr][J] - mean4[rr]));
                if (e < k) sa += (vv - mean4[rr]) * (vv - mean4[rr]);
              }
              sb += xbe * xbe;
              if (in) p.xa_out[o * k + e] = vv;
            }
#pragma unroll
            for (int o2 = 1; o2 < 16; o2 <<= 1) {
              sb += __shfl_xor(sb, o2, 64);
              sa += __shfl_xor(sa, o2, 64);
            }
            if (i16 == 0) {
              spart[1][wv][vi] = sb;
              spart[2][wv][vi] = sa;
            }
          }
          __syncthreads();
          if (wv == 0 && i16 == 0) {
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
              if (sidx4[rr] < 0) continue;
              const int vi = out_row(rr);
              const double sb = (spart[1][0][vi] + spart[1][1][vi]) + (spart[1][2][vi] + spart[1][3][vi]);
              const double sa = (spart[2][0][vi] + spart[2][1][vi]) + (spart[2][2][vi] + spart[2][3][vi]);
              const long long so = spread_index(p, lo, la, sidx4[rr]);
              p.spread_b[so] = sqrt(sb / (double)(float)k);
              p.spread_a[so] = sqrt(sa / (double)(float)k);
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k = 65..128: the matrix does not fit a wavefront's registers, so the two
// per-matrix phases keep it in LDS (one 128-thread workgroup per matrix,
// thread r owns row r); the QL recurrence is k_ql_log<128>.
// ---------------------------------------------------------------------------

// Householder tridiagonalisation with A in LDS, one thread per row (one
// workgroup of max(KMAX, 64) threads per matrix); same outputs as k_tridiag.
template <int KMAX>
__global__ __launch_bounds__(KMAX < 64 ? 64 : KMAX) void k_tridiag_lds(EigParams p) {
  constexpr int NT = KMAX < 64 ? 64 : KMAX;
  constexpr int NW = NT / 64;
  constexpr int LD = KMAX + 1;
  __shared__ double sA[KMAX * LD];
  __shared__ double sv[KMAX], sw[KMAX];
  __shared__ double se_s[KMAX + 1], st_s[KMAX + 1];
  __shared__ double red[2 * NW];
  const int r = threadIdx.x;
  const long long a = p.a0 + blockIdx.x;
  if (a >= p.a0 + p.n_an) return;
  const long long al = a - p.a0;
  const int k = p.k, kp = p.kp;
  const double* __restrict__ Ag = p.A + al * (long long)kp * kp;
  double* __restrict__ refl = p.refl + al * (long long)kp * kp;
  const bool row = r < KMAX;             // thread owns a row of the LDS matrix
  const bool in_r = r < k;
  if (row)
    for (int c = 0; c < KMAX; ++c)
      sA[c * LD + r] = (in_r && c < k) ? a_entry(Ag, p.a_f32, kp, r, c) : 0.0;
  int slot = 0;
  __syncthreads();
  for (int j = 0; j + 2 < k; ++j) {
    const bool act = in_r && r > j;
    const double x = act ? sA[j * LD + r] : 0.0;
    const double x0 = uniform(sA[j * LD + j + 1]);
    const double rest = block_sum<NW>((act && r > j + 1) ? x * x : 0.0, red, slot);
    if (rest == 0.0) {
      if (row) refl[(long long)j * kp + r] = 0.0;
      if (r == 0) { se_s[j] = x0; st_s[j] = 0.0; }
      continue;
    }
    const double nrm = sqrt(rest + x0 * x0);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;
    const double v0 = x0 - alpha;
    const double tau = 2.0 / (rest + v0 * v0);
    const double v = act ? ((r == j + 1) ? v0 : x) : 0.0;
    if (row) sv[r] = v;
    __syncthreads();
    double pr0 = 0.0, pr1 = 0.0;
    if (act) {
      int c = j + 1;
      for (; c + 1 < k; c += 2) {
        pr0 = fma(sA[c * LD + r], sv[c], pr0);
        pr1 = fma(sA[(c + 1) * LD + r], sv[c + 1], pr1);
      }
      if (c < k) pr0 = fma(sA[c * LD + r], sv[c], pr0);
    }
    const double pr = act ? (pr0 + pr1) * tau : 0.0;
    const double K = 0.5 * tau * block_sum<NW>(v * pr, red, slot);
    const double wr = act ? pr - K * v : 0.0;
    if (row) sw[r] = wr;
    __syncthreads();
    if (act)
      for (int c = j + 1; c < k; ++c) sA[c * LD + r] -= v * sw[c] + wr * sv[c];
    if (row) refl[(long long)j * kp + r] = v;
    if (r == 0) { se_s[j] = alpha; st_s[j] = tau; }
    __syncthreads();
  }
  __syncthreads();
  if (!row) return;
  double* __restrict__ dg = p.d + al * kp;
  double* __restrict__ eg = p.e + al * kp;
  double* __restrict__ tg = p.tau + al * kp;
  dg[r] = in_r ? sA[r * LD + r] : 0.0;
  double ev = 0.0, tv = 0.0;
  if (r + 2 < k) { ev = se_s[r]; tv = st_s[r]; }
  else if (r == k - 2) ev = sA[(k - 2) * LD + k - 1];
  eg[r] = ev;
  tg[r] = tv;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// true when the batch updates a state (not only transforms)
static bool has_state(const EigParams& p) { return p.xb || p.xb32; }

// the update kernel for the batch's width, state precision and output mode
template <int KMAX>
static void launch_update(hipStream_t st, const EigParams& p, unsigned nb) {
  if constexpr (KMAX == 128) {
    if (p.state_f32) {
      if (p.epi) hipLaunchKernelGGL((k_update128<float, 1>), dim3(nb), dim3(256), 0, st, p);
      else hipLaunchKernelGGL((k_update128<float, 0>), dim3(nb), dim3(256), 0, st, p);
    } else {
      if (p.epi) hipLaunchKernelGGL((k_update128<double, 1>), dim3(nb), dim3(256), 0, st, p);
      else hipLaunchKernelGGL((k_update128<double, 0>), dim3(nb), dim3(256), 0, st, p);
    }
  } else {
    if (p.state_f32) {
      if (p.epi) hipLaunchKernelGGL((k_update<KMAX, float, 1>), dim3(nb), dim3(256), 0, st, p);
      else hipLaunchKernelGGL((k_update<KMAX, float, 0>), dim3(nb), dim3(256), 0, st, p);
    } else {
      if (p.epi) hipLaunchKernelGGL((k_update<KMAX, double, 1>), dim3(nb), dim3(256), 0, st, p);
      else hipLaunchKernelGGL((k_update<KMAX, double, 0>), dim3(nb), dim3(256), 0, st, p);
    }
  }
}

hipError_t launch_eig(hipStream_t st, const EigParams& p, hipEvent_t* ev) {
  if (p.n_an <= 0) return hipSuccess;
  const unsigned nb = (unsigned)p.n_an;
  const unsigned nq = (unsigned)((p.n_an + 63) / 64);
  auto mark = [&](int i) -> hipError_t { return ev ? hipEventRecord(ev[i], st) : hipSuccess; };
  hipError_t e = mark(0);
  if (e != hipSuccess) return e;
  if (p.k <= 32) {
    if (p.tridiag_lds) hipLaunchKernelGGL(k_tridiag_lds<32>, dim3(nb), dim3(64), 0, st, p);
    else hipLaunchKernelGGL(k_tridiag<32>, dim3(nb), dim3(64), 0, st, p);
    if ((e = mark(1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_ql_log<32>, dim3(nq), dim3(64), 0, st, p);
    if ((e = mark(2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_vec<32>, dim3(nb), dim3(64), 0, st, p);
    if ((e = mark(3)) != hipSuccess) return e;
    if (has_state(p)) launch_update<32>(st, p, nb);
  } else if (p.k <= 64) {
    if (p.tridiag_lds) hipLaunchKernelGGL(k_tridiag_lds<64>, dim3(nb), dim3(64), 0, st, p);
    else hipLaunchKernelGGL(k_tridiag<64>, dim3(nb), dim3(64), 0, st, p);
    if ((e = mark(1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_ql_log<64>, dim3(nq), dim3(64), 0, st, p);
    if ((e = mark(2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_vec<64>, dim3(nb), dim3(64), 0, st, p);
    if ((e = mark(3)) != hipSuccess) return e;
    if (has_state(p)) launch_update<64>(st, p, nb);
  } else if (p.k <= 128) {
    if (p.tridiag_lds) hipLaunchKernelGGL(k_tridiag_lds<128>, dim3(nb), dim3(128), 0, st, p);
    else hipLaunchKernelGGL(k_tridiag128, dim3(nb), dim3(256), 0, st, p);
    if ((e = mark(1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_ql_log<128>, dim3(nq), dim3(64), 0, st, p);
    if ((e = mark(2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_vec128, dim3(nb), dim3(256), 0, st, p);
    if ((e = mark(3)) != hipSuccess) return e;
    if (has_state(p)) launch_update<128>(st, p, nb);
  } else {
    return hipErrorInvalidValue;
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return mark(4);
}

// analyses [off, off + n) of the batch p: batch-local arrays offset, global
// ones (status, states, trans_out, w_out) unchanged
static EigParams eig_slice(const EigParams& p, long long off, long long n) {
  EigParams q = p;
  const long long kp = p.kp;
  q.a0 = p.a0 + off;
  q.n_an = n;
  q.A = p.A + off * kp * kp;
  q.b = p.b + off * kp;
  q.refl = p.refl + off * kp * kp;
  q.tau = p.tau + off * kp;
  q.d = p.d + off * kp;
  q.e = p.e + off * kp;
  q.lam = p.lam + off * kp;
  q.rot = p.rot + off * (long long)p.rot_cap;
  q.hdr = p.hdr + off * (long long)p.hdr_cap;
  q.nchase = p.nchase + off;
  return q;
}

// one of the four split-eigensolver kernels (tridiagonalisation, QL,
// eigenvectors + trans, update) over the batch p
template <int KMAX>
static void launch_eig_stage_k(hipStream_t st, const EigParams& p, int stage) {
  const unsigned nb = (unsigned)p.n_an;
  const unsigned nq = (unsigned)((p.n_an + 63) / 64);
  if (stage == 0) {
    if (p.tridiag_lds) hipLaunchKernelGGL(k_tridiag_lds<KMAX>, dim3(nb), dim3(KMAX < 64 ? 64 : KMAX), 0, st, p);
    else if constexpr (KMAX == 128) hipLaunchKernelGGL(k_tridiag128, dim3(nb), dim3(256), 0, st, p);
    else hipLaunchKernelGGL(k_tridiag<KMAX>, dim3(nb), dim3(64), 0, st, p);
  } else if (stage == 1) {
    hipLaunchKernelGGL(k_ql_log<KMAX>, dim3(nq), dim3(64), 0, st, p);
  } else if (stage == 2) {
    if constexpr (KMAX == 128) hipLaunchKernelGGL(k_vec128, dim3(nb), dim3(256), 0, st, p);
    else hipLaunchKernelGGL(k_vec<KMAX>, dim3(nb), dim3(64), 0, st, p);
  } else if (has_state(p)) {
    launch_update<KMAX>(st, p, nb);
  }
}

static void launch_eig_stage(hipStream_t st, const EigParams& p, int stage) {
  if (p.k <= 32) launch_eig_stage_k<32>(st, p, stage);
  else if (p.k <= 64) launch_eig_stage_k<64>(st, p, stage);
  else launch_eig_stage_k<128>(st, p, stage);
}

hipError_t launch_zero_ints(hipStream_t st, int* p, int n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

// The list pass: QL, eigenvectors + trans and update over the analyses the
// first pass flagged (p.list, *p.list_n of them, at most p.list_cap).  Grids
// are sized for list_cap; workgroups past the count exit at once, so the
// launch needs no count on the host.
hipError_t launch_eig_rerun(hipStream_t st, const EigParams& p) {
  if (!p.list || p.list_cap <= 0) return hipSuccess;
  const unsigned nb = (unsigned)p.list_cap;
  const unsigned nq = (unsigned)((p.list_cap + 63) / 64);
  for (int stage = 1; stage < 4; ++stage) {
    if (p.k <= 32) {
      if (stage == 1) hipLaunchKernelGGL(k_ql_log<32>, dim3(nq), dim3(64), 0, st, p);
      else if (stage == 2) hipLaunchKernelGGL(k_vec<32>, dim3(nb), dim3(64), 0, st, p);
      else if (has_state(p)) launch_update<32>(st, p, nb);
    } else if (p.k <= 64) {
      if (stage == 1) hipLaunchKernelGGL(k_ql_log<64>, dim3(nq), dim3(64), 0, st, p);
      else if (stage == 2) hipLaunchKernelGGL(k_vec<64>, dim3(nb), dim3(64), 0, st, p);
      else if (has_state(p)) launch_update<64>(st, p, nb);
    } else {
      if (stage == 1) hipLaunchKernelGGL(k_ql_log<128>, dim3(nq), dim3(64), 0, st, p);
      else if (stage == 2) hipLaunchKernelGGL(k_vec128, dim3(nb), dim3(256), 0, st, p);
      else if (has_state(p)) launch_update<128>(st, p, nb);
    }
  }
  return hipGetLastError();
}

hipError_t launch_eig_pipelined(hipStream_t st, hipStream_t st2, hipEvent_t e_first,
                                hipEvent_t e_join, const EigParams& p, hipEvent_t* ev) {
  if (p.k > 128 || p.n_an < 2 * kEigPipeMinHalf) return launch_eig(st, p, ev);
  // the QL recurrence holds a CU's LDS with one wavefront and leaves its other
  // SIMDs idle; the second half of the batch starts on st2 once the first
  // half's tridiagonalisation is done, so its kernels run beside the first
  // half's (tridiagonalisation beside QL, QL beside the eigenvectors)
  const long long n1 = (p.n_an + 1) / 2;
  const EigParams pa = eig_slice(p, 0, n1), pb = eig_slice(p, n1, p.n_an - n1);
  auto mark = [&](int i) -> hipError_t { return ev ? hipEventRecord(ev[i], st) : hipSuccess; };
  hipError_t e = mark(0);
  if (e != hipSuccess) return e;
  launch_eig_stage(st, pa, 0);
  if ((e = hipEventRecord(e_first, st)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(st2, e_first, 0)) != hipSuccess) return e;
  for (int stage = 0; stage < 4; ++stage) launch_eig_stage(st2, pb, stage);
  if ((e = hipEventRecord(e_join, st2)) != hipSuccess) return e;
  if ((e = mark(1)) != hipSuccess) return e;
  launch_eig_stage(st, pa, 1);
  if ((e = mark(2)) != hipSuccess) return e;
  launch_eig_stage(st, pa, 2);
  if ((e = mark(3)) != hipSuccess) return e;
  launch_eig_stage(st, pa, 3);
  if ((e = hipStreamWaitEvent(st, e_join, 0)) != hipSuccess) return e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return mark(4);
}

}  // namespace dida
