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
//   k_apply     one wavefront per matrix, lane r keeps row r of Q in VGPRs:
//               Q = H_0 ... H_{k-3} by forward accumulation, the logged
//               rotations applied in order (V = Q Z), then
//               overline_w = V diag(1/lambda) V^T b,
//               T = V diag(sqrt((k-1)/lambda)) V^T, trans = T + overline_w 1^T
//               (letkf_core.F90:129-225) and the ensemble update
//               (xb_2_xa, letkf_compute.F90:259-329, or the per-level update
//               of the vertical loop, letkf_compute.F90:501-552).
//
// Matrices whose QL did not converge or whose rotation log overflowed are
// flagged (status != 0) and skipped by k_apply; the host reruns them with the
// single-kernel transform.
#include "letkf_device.h"
#include "letkf_kernels.h"

namespace dida {

// ---------------------------------------------------------------------------
// k_tridiag
// ---------------------------------------------------------------------------
template <int KMAX>
__global__ __launch_bounds__(64) void k_tridiag(EigParams p) {
  constexpr int NB = KMAX / 8;
  __shared__ double sv[KMAX], sw[KMAX];
  __shared__ double se_s[KMAX + 1], st_s[KMAX + 1];
  const int lane = threadIdx.x;
  const long long a = p.a0 + blockIdx.x;
  if (a >= p.a0 + p.n_an) return;
  const long long al = a - p.a0;
  const int k = p.k, kp = p.kp;
  const double* __restrict__ Ag = p.A + al * (long long)kp * kp;
  double* __restrict__ refl = p.refl + al * (long long)kp * kp;
  const bool in_r = lane < k;

  double arow[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c)
    arow[c] = (in_r && c < k) ? Ag[(long long)c * kp + lane] : 0.0;
  double xcol = arow[0];

  for (int j = 0; j + 2 < k; ++j) {
    const bool act = in_r && lane > j;
    const double x = act ? xcol : 0.0;
    const double x0 = __shfl(xcol, j + 1, 64);
    const double rest = wave_sum((act && lane > j + 1) ? x * x : 0.0);
    double xn = 0.0;
    if (rest == 0.0) {
      // column already in tridiagonal form
#pragma unroll
      for (int c = 0; c < KMAX; ++c) xn = (c == j + 1) ? arow[c] : xn;
      if (lane < KMAX) refl[(long long)j * kp + lane] = 0.0;
      if (lane == 0) { se_s[j] = x0; st_s[j] = 0.0; }
      xcol = xn;
      continue;
    }
    const double nrm = sqrt(rest + x0 * x0);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;
    const double v0 = x0 - alpha;
    const double tau = 2.0 / (rest + v0 * v0);
    const double v = act ? ((lane == j + 1) ? v0 : x) : 0.0;
    if (lane < KMAX) sv[lane] = v;
    __syncthreads();
    // p = tau A v over the trailing columns (sv is zero for c <= j)
    double pr = 0.0;
#pragma unroll
    for (int B = 0; B < NB; ++B) {
      if (8 * B + 7 <= j) continue;
#pragma unroll
      for (int t = 0; t < 8; ++t) pr = fma(arow[8 * B + t], sv[8 * B + t], pr);
    }
    pr = act ? pr * tau : 0.0;
    const double K = 0.5 * tau * wave_sum(v * pr);
    const double wr = act ? pr - K * v : 0.0;
    if (lane < KMAX) sw[lane] = wr;
    __syncthreads();
#pragma unroll
    for (int B = 0; B < NB; ++B) {
      if (8 * B + 7 <= j) continue;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int c = 8 * B + t;
        arow[c] -= v * sw[c] + wr * sv[c];
        xn = (c == j + 1) ? arow[c] : xn;
      }
    }
    if (lane < KMAX) refl[(long long)j * kp + lane] = v;
    if (lane == 0) { se_s[j] = alpha; st_s[j] = tau; }
    xcol = xn;
    __syncthreads();
  }
  // diagonal and the last subdiagonal entry
  double dr = 0.0, ek2 = 0.0;
#pragma unroll
  for (int c = 0; c < KMAX; ++c) {
    dr = (c == lane) ? arow[c] : dr;
    ek2 = (c == k - 2) ? arow[c] : ek2;
  }
  ek2 = __shfl(ek2, k - 1, 64);
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
template <int KMAX>
__global__ __launch_bounds__(64) void k_ql_log(EigParams p) {
  __shared__ double sd[KMAX][64];
  __shared__ double se[KMAX][64];
  const int lane = threadIdx.x;
  const long long a = p.a0 + (long long)blockIdx.x * 64 + lane;
  const bool valid = a < p.a0 + p.n_an;
  const long long al = a - p.a0;
  const int k = p.k, kp = p.kp;
  double tnorm = 0.0;
  for (int i = 0; i < KMAX; ++i) {
    double dv = 0.0, ev = 0.0;
    if (valid && i < k) {
      dv = p.d[al * kp + i];
      ev = p.e[al * kp + i];
    }
    sd[i][lane] = dv;
    se[i][lane] = ev;
    tnorm = fmax(tnorm, fabs(dv) + fabs(ev));
  }
  if (!valid) return;
  const double eps = 2.220446049250313e-16;
  double2* __restrict__ rot = p.rot + al * (long long)p.rot_cap;
  int4* __restrict__ hdr = p.hdr + al * (long long)p.hdr_cap;
  int nrot = 0, nhdr = 0, status = 0;
  for (int l = 0; l < k && status == 0; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < k - 1; ++m) {
        const double dd = fabs(sd[m][lane]) + fabs(sd[m + 1][lane]);
        const double em = fabs(se[m][lane]);
        if (em <= eps * dd || em <= eps * tnorm) break;
      }
      if (m == l) break;
      if (++iter > 30) { status = l + 1; break; }
      if (nhdr >= p.hdr_cap || nrot + (m - l) > p.rot_cap) { status = -1; break; }
      double g = (sd[l + 1][lane] - sd[l][lane]) / (2.0 * se[l][lane]);
      double r = hypot(g, 1.0);
      g = sd[m][lane] - sd[l][lane] + se[l][lane] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      const int off = nrot;
      int i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        const double ei = se[i][lane];
        const double f = s * ei, bb = c * ei;
        r = hypot(f, g);
        se[i + 1][lane] = r;
        if (r == 0.0) {
          sd[i + 1][lane] -= pp;
          se[m][lane] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = sd[i + 1][lane] - pp;
        r = (sd[i][lane] - g) * s + 2.0 * c * bb;
        pp = s * r;
        sd[i + 1][lane] = g + pp;
        g = c * r - bb;
        rot[nrot++] = make_double2(c, s);
      }
      const int lo = underflow ? i + 1 : l;
      if (m - lo > 0) hdr[nhdr++] = make_int4(lo, m, off, 0);
      if (underflow) continue;
      sd[l][lane] -= pp;
      se[l][lane] = g;
      se[m][lane] = 0.0;
    }
  }
  for (int i = 0; i < k; ++i) p.lam[al * kp + i] = sd[i][lane];
  p.nchase[al] = nhdr;
  p.status[a] = status;
}

// ---------------------------------------------------------------------------
// k_apply
// ---------------------------------------------------------------------------
template <int KMAX>
__global__ __launch_bounds__(64) void k_apply(EigParams p) {
  constexpr int NB = KMAX / 8;
  constexpr int TT = KMAX / 16;          // 16x16 tiles per dimension
  constexpr int LDV = KMAX + 1;
  __shared__ double sV[KMAX * LDV];
  __shared__ double sf[KMAX], sg[KMAX], sb[KMAX], st[KMAX], sw[KMAX];
  const int lane = threadIdx.x;
  const long long a = p.a0 + blockIdx.x;
  if (a >= p.a0 + p.n_an) return;
  const long long al = a - p.a0;
  if (p.status[a] != 0) return;          // rerun by the single-kernel transform
  const int k = p.k, kp = p.kp;
  const bool in_r = lane < k;

  // ---- Q rows by forward accumulation: Z <- Z H_j ----
  double z[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) z[c] = (c == lane) ? 1.0 : 0.0;
  const double* __restrict__ refl = p.refl + al * (long long)kp * kp;
  const double* __restrict__ tg = p.tau + al * kp;
  for (int j = 0; j + 2 < k; ++j) {
    const double tau = tg[j];
    if (tau == 0.0) continue;
    const double* __restrict__ vj = refl + (long long)j * kp;
    double dot = 0.0;
#pragma unroll
    for (int B = 0; B < NB; ++B) {
      if (8 * B + 7 <= j) continue;
#pragma unroll
      for (int t = 0; t < 8; ++t) dot = fma(z[8 * B + t], vj[8 * B + t], dot);
    }
    dot *= tau;
#pragma unroll
    for (int B = 0; B < NB; ++B) {
      if (8 * B + 7 <= j) continue;
#pragma unroll
      for (int t = 0; t < 8; ++t) z[8 * B + t] = fma(-dot, vj[8 * B + t], z[8 * B + t]);
    }
  }

  // ---- the logged QL rotations, chase by chase ----
  const double2* __restrict__ rot = p.rot + al * (long long)p.rot_cap;
  const int4* __restrict__ hdr = p.hdr + al * (long long)p.hdr_cap;
  const int nch = p.nchase[al];
  for (int ch = 0; ch < nch; ++ch) {
    const int4 h = hdr[ch];
    const int lo = h.x, hi = h.y, off = h.z;
    const int bhi = (hi - 1) >> 3, blo = lo >> 3;
    double carry = 0.0;
#pragma unroll
    for (int B = NB - 1; B >= 0; --B) {
      if (B > bhi || B < blo) continue;
      const int top = (8 * B + 7 < KMAX - 1) ? 8 * B + 7 : KMAX - 2;
      if (B == bhi) carry = z[top + 1];
#pragma unroll
      for (int t = 7; t >= 0; --t) {
        const int i = 8 * B + t;
        if (i > KMAX - 2) continue;
        double c = 1.0, s = 0.0;
        if (i >= lo && i < hi) {
          const double2 cs = rot[off + (hi - 1 - i)];
          c = cs.x;
          s = cs.y;
        }
        const double zi = z[i];
        z[i + 1] = s * zi + c * carry;
        carry = c * zi - s * carry;
      }
      if (B == blo) z[8 * B] = carry;
    }
  }

  // ---- eigenvalue functions (letkf_core.F90:107-113, guard for reduced modes) ----
  double lam = in_r ? p.lam[al * kp + lane] : 0.0;
  if (p.guard) {
    const double mn = wave_min(in_r ? lam : 1e300);
    if (mn < 0.0) lam = lam - mn;
  }
  lam = lam * 1.0 + (double)(k - 1) / p.inflation;
  if (lane < KMAX) {
    sf[lane] = in_r ? sqrt((double)(k - 1) / lam) : 0.0;
    sg[lane] = in_r ? 1.0 / lam : 0.0;
    sb[lane] = in_r ? p.b[al * kp + lane] : 0.0;
  }
  // V rows to LDS (padding rows/cols beyond k are the identity, harmless:
  // f, g and b are zero there)
  if (lane < KMAX) {
#pragma unroll
    for (int c = 0; c < KMAX; ++c) sV[lane * LDV + c] = z[c];
  }
  __syncthreads();
  // overline_w = V diag(1/lambda) V^T b  (letkf_core.F90:129-195 regrouped)
  if (lane < KMAX) {
    double tb = 0.0;
    for (int r = 0; r < KMAX; ++r) tb = fma(sV[r * LDV + lane], sb[r], tb);
    st[lane] = tb * sg[lane];
  }
  __syncthreads();
  double wbar = 0.0;
#pragma unroll
  for (int c = 0; c < KMAX; ++c) wbar = fma(z[c], st[c], wbar);
  if (lane < KMAX) sw[lane] = in_r ? wbar : 0.0;
  __syncthreads();

  // ---- T = V diag(f) V^T on f64 MFMA (16x16x4), all TT x TT tiles ----
  // A(I,s) = V[16I + l%16][4s + l/16], B(s,J) = f_{4s+l/16} V[16J + l%16][4s + l/16];
  // tile (I,J) reg r of lane l holds T[16I + l/16 + 4r][16J + l%16], which is
  // also the B operand of the update GEMM for k-step s = 4I + r.
  const int i16 = lane & 15, q = lane >> 4;
  d4 acc[TT][TT];
#pragma unroll
  for (int I = 0; I < TT; ++I)
#pragma unroll
    for (int J = 0; J < TT; ++J) acc[I][J] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
  for (int sK = 0; sK < KMAX / 4; ++sK) {
    const int c = 4 * sK + q;
    const double fc = sf[c];
    double fr[TT];
#pragma unroll
    for (int I = 0; I < TT; ++I) fr[I] = sV[(16 * I + i16) * LDV + c];
#pragma unroll
    for (int I = 0; I < TT; ++I)
#pragma unroll
      for (int J = 0; J < TT; ++J)
        acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[I], fc * fr[J], acc[I][J], 0, 0, 0);
  }
  // trans(f, e) = T(f, e) + overline_w(f)  (letkf_core.F90:221-225)
#pragma unroll
  for (int I = 0; I < TT; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double wf = sw[16 * I + q + 4 * r];
#pragma unroll
      for (int J = 0; J < TT; ++J) acc[I][J][r] += wf;
    }
  if (p.trans_out) {
    double* tr = p.trans_out + a * (long long)k * k;
#pragma unroll
    for (int I = 0; I < TT; ++I)
#pragma unroll
      for (int J = 0; J < TT; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int f = 16 * I + q + 4 * r, e = 16 * J + i16;
          if (f < k && e < k) tr[(long long)e * k + f] = acc[I][J][r];
        }
  }
  if (p.w_out && in_r) p.w_out[a * (long long)k + lane] = wbar;
  if (!p.xb) return;

  // ---- ensemble update on f64 MFMA: Xa(16 vectors x k) = Xb * trans ----
  const int ah = (int)(a / p.nzc), kc = (int)(a % p.nzc);
  const int ic = ah % p.g.nlat_c, jc = ah / p.g.nlat_c;
  const int sj = ic * p.g.cube_n + 1, ej = min(sj + p.g.cube_n - 1, p.g.nlat);
  const int si = jc * p.g.cube_n + 1, ei = min(si + p.g.cube_n - 1, p.g.nlon);
  const long long lat_ext = p.g.nlat + 2 * p.halo;
  int sk = 1, ek = p.nz;
  if (p.vertical) {
    sk = kc * p.cube_nz + 1;
    ek = min(sk + p.cube_nz - 1, p.nz);
  }
  // runs of contiguous state vectors of one column: (state start, length)
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
  for (int lo = si; lo <= ei; ++lo) {
    for (int la = sj; la <= ej; ++la) {
      const long long off = ((long long)(lo - 1 + p.halo) * lat_ext + (la - 1 + p.halo)) * p.n;
      for (int v0 = 0; v0 < nvec; v0 += 16) {
        // state index of vector v0 + i for i = 0..15 (lane i16 maps its own)
        auto sidx = [&](int v) {
          int rem = v;
          for (int t = 0; t < nrun; ++t) {
            if (rem < rl[t]) return rs[t] + rem;
            rem -= rl[t];
          }
          return -1;
        };
        // stage 16 xb vectors (k members each) in LDS, coalesced per vector
        __syncthreads();
        for (int u = 0; u < 16; ++u) {
          const int st_i = (v0 + u < nvec) ? sidx(v0 + u) : -1;
          if (lane < KMAX)
            sV[u * LDV + lane] = (st_i >= 0 && in_r) ? p.xb[(off + st_i) * k + lane] : 0.0;
        }
        __syncthreads();
        d4 out[TT];
#pragma unroll
        for (int J = 0; J < TT; ++J) out[J] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int sK = 0; sK < KMAX / 4; ++sK) {
          const double xa_frag = sV[i16 * LDV + 4 * sK + q];
#pragma unroll
          for (int J = 0; J < TT; ++J)
            out[J] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa_frag, acc[sK >> 2][J][sK & 3],
                                                          out[J], 0, 0, 0);
        }
        // lane holds Xa[vector q + 4r][member 16J + i16]
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int v = v0 + q + 4 * r;
          const int st_i = (v < nvec) ? sidx(v) : -1;
          const double ab = (st_i >= 0) ? p.axb[off + st_i] : 0.0;
          double sum = 0.0;
          double xe[TT];
#pragma unroll
          for (int J = 0; J < TT; ++J) {
            const int e = 16 * J + i16;
            xe[J] = (e < k) ? out[J][r] + ab : 0.0;
            sum += xe[J];
          }
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) sum += __shfl_xor(sum, o, 64);
          const double mean = sum / k;
          if (st_i >= 0) {
            const long long o = off + st_i;
#pragma unroll
            for (int J = 0; J < TT; ++J) {
              const int e = 16 * J + i16;
              if (e < k) p.xa[o * k + e] = xe[J] - mean;
            }
            if (i16 == 0) p.axa[o] = mean;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_eig(hipStream_t st, const EigParams& p) {
  if (p.n_an <= 0) return hipSuccess;
  const unsigned nb = (unsigned)p.n_an;
  const unsigned nq = (unsigned)((p.n_an + 63) / 64);
  if (p.k <= 32) {
    hipLaunchKernelGGL(k_tridiag<32>, dim3(nb), dim3(64), 0, st, p);
    hipLaunchKernelGGL(k_ql_log<32>, dim3(nq), dim3(64), 0, st, p);
    hipLaunchKernelGGL(k_apply<32>, dim3(nb), dim3(64), 0, st, p);
  } else if (p.k <= 64) {
    hipLaunchKernelGGL(k_tridiag<64>, dim3(nb), dim3(64), 0, st, p);
    hipLaunchKernelGGL(k_ql_log<64>, dim3(nq), dim3(64), 0, st, p);
    hipLaunchKernelGGL(k_apply<64>, dim3(nb), dim3(64), 0, st, p);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dida
