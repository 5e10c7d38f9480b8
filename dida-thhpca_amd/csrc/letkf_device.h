// letkf_device.h -- device-side helpers shared by the LETKF kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dida {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

// Raw buffer loads of 8 and 16 bytes (buffer_load_dwordx2 / dwordx4 offen;
// an offset past the resource's num_records bytes reads 0), bound to the
// LLVM intrinsics: this compiler lowers __builtin_amdgcn_raw_buffer_load_b128
// to a single-dword load.
__device__ f2 raw_buffer_load_f2(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ f4 raw_buffer_load_f4(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");


// pro_info.F90:148-161: pi = atan(1.0)*4 evaluated in real(4), and the
// real(4) literal a = 6371.393, both promoted to real(8).
constexpr double kPiF32 = 3.1415927410125732;
constexpr double kEarthF32 = 6371.39306640625;
// letkf_yb.F90:273-278 real(4) literals 1.666667 and 0.6666667
constexpr double kC1666667 = 1.66666698455810546875;
constexpr double kC0666667 = 0.666666686534881591796875;

// Gaspari-Cohn, letkf_yb.F90:262-286, in the Fortran operation order
// without contraction.
__device__ __forceinline__ double schr2(double ri) {
#pragma clang fp contract(off)
  double rci;
  if (ri > 2.0) {
    rci = 0.0;
  } else if (ri > 1.0) {
    double rj = ri / 12.0 - 0.5;
    rj = rj * ri + 0.625;
    rj = rj * ri + kC1666667;
    rci = (rj * ri - 5.0) * ri + 4.0 - kC0666667 / ri;
  } else if (ri > 0.0) {
    double rk = ri * ri;
    double rj = (-0.25 * ri + 0.5) * ri + 0.625;
    rci = (rj * ri - kC1666667) * rk + 1.0;
  } else {
    rci = 1.0;
  }
  return rci;
}

// letkf_yb.F90:287-302
__device__ __forceinline__ double schr(double ri) { return ri > 2.0 ? 0.0 : 1.0; }

__device__ __forceinline__ double gc_weight(double ri, int gc_filter) {
  return gc_filter ? schr2(ri) : schr(ri);
}

// latlon_dist_r8 with the per-point factors precomputed:
// c1 = cos(x1/180*pi), s1 = sin(x1/180*pi), l1 = y1/180*pi (same for 2).
__device__ __forceinline__ double latlon_dist_pre(double c1, double s1, double l1, double c2,
                                                  double s2, double l2) {
#pragma clang fp contract(off)
  double c3 = cos(l1 - l2);
  double v = c1 * c2 * c3 + s1 * s2;
  v = (v > 1.0) ? 1.0 : v;   // min(v, 1.0); NaN passes through like Fortran MIN
  return kEarthF32 * acos(v);
}

__device__ __forceinline__ double deg2rad_ref(double x) {
#pragma clang fp contract(off)
  return x / 180.0 * kPiF32;
}

// A value every active lane holds identically, moved to a scalar register:
// branches and loop bounds on it then compile to uniform (scalar) control
// flow instead of exec-mask divergence.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffu));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Butterfly sums: every lane ends with the bitwise-same value (IEEE addition
// is commutative, so partners add the same two operands).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Cross-lane moves of a double on the VALU: a DPP permutation within rows of
// 16 lanes, and a read of one lane (the lane index must be wavefront-uniform).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Reductions for code where all 64 lanes are active: four DPP steps reduce
// each row of 16 lanes (xor 1, xor 2, then the half-row and row mirrors, which
// pair lanes holding equal partial values) on the VALU, then two butterfly
// shuffles (xor 16, xor 32) combine the rows.  Every lane gets the bitwise
// same value (partners add the same two operands).
template <typename Op>
__device__ __forceinline__ double wave_reduce_full(double v, Op op) {
  v = op(v, dpp_d<0xB1>(v));    // quad_perm [1,0,3,2]
  v = op(v, dpp_d<0x4E>(v));    // quad_perm [2,3,0,1]
  v = op(v, dpp_d<0x141>(v));   // row_half_mirror
  v = op(v, dpp_d<0x140>(v));   // row_mirror
  v = op(v, __shfl_xor(v, 16, 64));
  return op(v, __shfl_xor(v, 32, 64));
}
__device__ __forceinline__ double wave_sum_full(double v) {
  return wave_reduce_full(v, [](double a, double b) { return a + b; });
}
__device__ __forceinline__ double wave_min_full(double v) {
  return wave_reduce_full(v, [](double a, double b) { return fmin(a, b); });
}

// exclusive prefix sum of an int over the 64 lanes
__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

}  // namespace dida
