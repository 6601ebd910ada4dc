// Cross-lane building blocks for np=4 spectral elements on CDNA (wave64).
//
// A 16-lane group ("slab") holds the 16 GLL points of one element at one
// level: lane l of the group is point pt = l & 15 = igp*4 + jgp, i.e. jgp is
// the fastest index (Fortran i), igp the slower one (Fortran j).
//   * row_get<K>(x): value at (igp, K) of the caller's slab  -> DPP quad_perm
//   * col_get<K>(x): value at (K, jgp) of the caller's slab  -> ds_swizzle
//   * slab_sum_serial(x): (((0 + x0) + x1) + x2) + ... + x15, left to right,
//     in every lane of the slab (64-bit DPP row_newbcast: each step adds
//     point j's value to every lane of the row) -- the same rounding as a
//     serial loop over the points.
//   * slab_min / slab_max: exact, any order.
// All operations act on doubles as two 32-bit halves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hx {

constexpr int NP = 4;
constexpr int NP2 = 16;
constexpr double RGAS = 287.04;
constexpr double CP = 1005.0;
constexpr double RWATER_VAPOR = 461.5;
constexpr double KAPPA = RGAS / CP;
constexpr double RREARTH = 1.0 / 6.376e6;

__device__ __forceinline__ int lo32(double x) { return __double2loint(x); }
__device__ __forceinline__ int hi32(double x) { return __double2hiint(x); }
__device__ __forceinline__ double mk64(int hi, int lo) { return __hiloint2double(hi, lo); }

// DPP moves without an "old" operand: with bound_ctrl set, a lane whose
// source is outside its row reads 0, so the destination needs no zeroing
// move first (quad_perm sources are always inside the quad).
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int x) { return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true); }

// DPP quad_perm broadcast of lane K of each quad (same igp row, jgp = K).
template <int K>
__device__ __forceinline__ double row_get(double x) {
  constexpr int ctrl = K * 0x55;  // quad_perm:[K,K,K,K]
  return mk64(dpp_mov<ctrl>(hi32(x)), dpp_mov<ctrl>(lo32(x)));
}

// ds_swizzle bit-mask mode within 32 lanes: src = (lane & 0x13) | (K << 2):
// keeps jgp (bits 0-1) and the half-group bit 4, forces igp = K.
template <int K>
__device__ __forceinline__ double col_get(double x) {
  constexpr int pattern = 0x13 | ((K << 2) << 5);
  const int l = __builtin_amdgcn_ds_swizzle(lo32(x), pattern);
  const int h = __builtin_amdgcn_ds_swizzle(hi32(x), pattern);
  return mk64(h, l);
}

// Serial (left-to-right over points 0..15) sum, result in every lane of the
// slab: s = 0, then s = fma(x_j, 1, s) for j = 0..15, where x_j is lane j of
// the 16-lane row broadcast by row_newbcast:j (the one DPP control 64-bit
// VALU operations take on CDNA3/4).  fma(x, 1, s) rounds once, exactly as
// s + x.  The leading s_nop covers the two wait states a DPP source needs
// after a VALU write; the trailing one covers the same for a DPP read of the
// result by code the compiler schedules after this block.
__device__ __forceinline__ double slab_sum_serial(double v) {
  double s;
  const double one = 1.0;
  asm volatile(
      "v_mov_b64 %0, 0\n\t"
      "s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:15 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1"
      : "=&v"(s)
      : "v"(v), "v"(one));
  return s;
}

// xor-butterfly within 16 lanes via ds_swizzle (and_mask 0x1F, xor_mask M)
template <int M>
__device__ __forceinline__ double xor_get(double x) {
  constexpr int pattern = 0x1F | (M << 10);
  const int l = __builtin_amdgcn_ds_swizzle(lo32(x), pattern);
  const int h = __builtin_amdgcn_ds_swizzle(hi32(x), pattern);
  return mk64(h, l);
}

// Division by a divisor shared by many quotients, the same value as a / b
// (bit for bit, except that a zero quotient may come back as +0): with y = 1/b correctly rounded, q0 = a*y is within one ulp of a/b, the
// remainder a - q0*b is exact in one FMA, and q0 + rem*y rounded once is the
// correctly rounded quotient (Markstein's theorem on FMA division, IBM J.
// Res. Dev. 34(1), 1990), for finite operands whose remainder does not
// underflow (|a| far above 2^-960).  Three VALU operations per quotient
// instead of the ten of an IEEE division.
struct Recip {
  double b, y;
};
__device__ __forceinline__ Recip recip(double b) { return {b, 1.0 / b}; }
__device__ __forceinline__ double div_r(double a, const Recip &r) {
  const double q0 = a * r.y;
  const double rem = __builtin_fma(-q0, r.b, a);
  return __builtin_fma(rem, r.y, q0);
}
// the same with a constant divisor (y folded by the compiler)
template <int B>
__device__ __forceinline__ double div_c(double a) {
  return div_r(a, Recip{(double)B, 1.0 / (double)B});
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// One v_min_f64 / v_max_f64: fmin/fmax of a swizzled value would first be
// quieted by a canonicalising v_max_f64 (the operands here are never NaN).
// Of two equal values of opposite sign either zero may come back; the
// value is the same.
__device__ __forceinline__ double vmin64(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double vmax64(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

__device__ __forceinline__ double slab_min(double x) {
  x = vmin64(x, xor_get<1>(x));
  x = vmin64(x, xor_get<2>(x));
  x = vmin64(x, xor_get<4>(x));
  x = vmin64(x, xor_get<8>(x));
  return x;
}
__device__ __forceinline__ double slab_max(double x) {
  x = vmax64(x, xor_get<1>(x));
  x = vmax64(x, xor_get<2>(x));
  x = vmax64(x, xor_get<4>(x));
  x = vmax64(x, xor_get<8>(x));
  return x;
}

// true if any active lane of the wave has pred set (wave-uniform)
__device__ __forceinline__ bool wave_any(bool pred) { return __ballot(pred) != 0; }

// ---------------------------------------------------------------- metric
// Per-lane constants of one element: the 2x2 tensors at the lane's point in
// reference C++ index order (D(idim,jdim) = flat[idim*2+jdim]), and the
// rows/columns of dvv the lane's contractions need.
struct DvvLane {
  double J[4];   // dvv[jgp][k]
  double I[4];   // dvv[igp][k]
  double TJ[4];  // dvv[k][jgp]
  double TI[4];  // dvv[k][igp]
};

__device__ __forceinline__ void load_dvv(const double *__restrict__ dvv, int pt, DvvLane &d) {
  const int ig = pt >> 2, jg = pt & 3;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    d.J[k] = dvv[jg * 4 + k];
    d.I[k] = dvv[ig * 4 + k];
    d.TJ[k] = dvv[k * 4 + jg];
    d.TI[k] = dvv[k * 4 + ig];
  }
}

// sum_k c[k] * row_get<k>(x) in k order, starting from 0
__device__ __forceinline__ double row_contract(const double c[4], double x) {
  double s = 0;
  s += c[0] * row_get<0>(x);
  s += c[1] * row_get<1>(x);
  s += c[2] * row_get<2>(x);
  s += c[3] * row_get<3>(x);
  return s;
}
__device__ __forceinline__ double col_contract(const double c[4], double x) {
  double s = 0;
  s += c[0] * col_get<0>(x);
  s += c[1] * col_get<1>(x);
  s += c[2] * col_get<2>(x);
  s += c[3] * col_get<3>(x);
  return s;
}

}  // namespace hx

// ------------------------------------------------------------------ FP32
// Single-precision counterparts for the mixed form: one DPP / swizzle per
// value, and dot products with compensated accumulation (exact product
// error from an FMA, exact sum error from TwoSum, summed separately and
// added at the end: the Dot2 scheme of Ogita, Rump and Oishi), which keeps a
// 4-term contraction accurate to a few FP32 ulps of its result.
namespace hx {

template <int K>
__device__ __forceinline__ float row_getf(float x) {
  return __int_as_float(dpp_mov<K * 0x55>(__float_as_int(x)));
}
template <int K>
__device__ __forceinline__ float col_getf(float x) {
  constexpr int pattern = 0x13 | ((K << 2) << 5);
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), pattern));
}

// s + e accumulates c*x exactly up to the final rounding
__device__ __forceinline__ void dot2_step(float c, float x, float &s, float &e) {
  const float p = c * x;
  const float pe = __builtin_fmaf(c, x, -p);  // c*x - p exactly
  const float t = s + p;
  const float z = t - s;
  const float te = (s - (t - z)) + (p - z);  // s + p - t exactly (TwoSum)
  s = t;
  e += pe + te;
}

__device__ __forceinline__ float row_contractf(const float c[4], float x) {
  float s = 0.f, e = 0.f;
  dot2_step(c[0], row_getf<0>(x), s, e);
  dot2_step(c[1], row_getf<1>(x), s, e);
  dot2_step(c[2], row_getf<2>(x), s, e);
  dot2_step(c[3], row_getf<3>(x), s, e);
  return s + e;
}
__device__ __forceinline__ float col_contractf(const float c[4], float x) {
  float s = 0.f, e = 0.f;
  dot2_step(c[0], col_getf<0>(x), s, e);
  dot2_step(c[1], col_getf<1>(x), s, e);
  dot2_step(c[2], col_getf<2>(x), s, e);
  dot2_step(c[3], col_getf<3>(x), s, e);
  return s + e;
}

struct DvvLaneF {
  float J[4], I[4], TJ[4], TI[4];
};
__device__ __forceinline__ void load_dvvf(const double *__restrict__ dvv, int pt, DvvLaneF &d) {
  const int ig = pt >> 2, jg = pt & 3;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    d.J[k] = (float)dvv[jg * 4 + k];
    d.I[k] = (float)dvv[ig * 4 + k];
    d.TJ[k] = (float)dvv[k * 4 + jg];
    d.TI[k] = (float)dvv[k * 4 + ig];
  }
}

}  // namespace hx
