// Cross-lane building blocks for np=4 spectral elements on CDNA (wave64).
//
// A 16-lane group ("slab") holds the 16 GLL points of one element at one
// level: lane l of the group is point pt = l & 15 = igp*4 + jgp, i.e. jgp is
// the fastest index (Fortran i), igp the slower one (Fortran j).
//   * row_get<K>(x): value at (igp, K) of the caller's slab  -> DPP quad_perm
//   * col_get<K>(x): value at (K, jgp) of the caller's slab  -> ds_swizzle
//   * slab_sum_serial(x): ((x0 + x1) + x2) + ... + x15, left to right, in
//     every lane of the slab (systolic DPP row_shr:1 chain, then broadcast
//     of lane 15) -- the same rounding as a serial loop over the points.
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

// value of lane 15 of each 16-lane row, broadcast to the row
__device__ __forceinline__ double row_last(double x) {
  constexpr int pattern = 0x10 | (0x0F << 5);
  const int l = __builtin_amdgcn_ds_swizzle(lo32(x), pattern);
  const int h = __builtin_amdgcn_ds_swizzle(hi32(x), pattern);
  return mk64(h, l);
}

// row_shr:1 within 16-lane rows, lane 0 of each row receives 0.0
__device__ __forceinline__ double shr1(double x) { return mk64(dpp_mov<0x111>(hi32(x)), dpp_mov<0x111>(lo32(x))); }

// Serial (left-to-right over points 0..15) sum, result in every lane of the slab.
__device__ __forceinline__ double slab_sum_serial(double v) {
  double t = v;
#pragma unroll
  for (int j = 0; j < 15; ++j) t = shr1(t) + v;
  return row_last(t);
}

// Two independent serial sums interleaved (more ILP).
__device__ __forceinline__ void slab_sum_serial2(double v, double w, double &sv, double &sw) {
  double t = v, u = w;
#pragma unroll
  for (int j = 0; j < 15; ++j) {
    t = shr1(t) + v;
    u = shr1(u) + w;
  }
  sv = row_last(t);
  sw = row_last(u);
}

// xor-butterfly within 16 lanes via ds_swizzle (and_mask 0x1F, xor_mask M)
template <int M>
__device__ __forceinline__ double xor_get(double x) {
  constexpr int pattern = 0x1F | (M << 10);
  const int l = __builtin_amdgcn_ds_swizzle(lo32(x), pattern);
  const int h = __builtin_amdgcn_ds_swizzle(hi32(x), pattern);
  return mk64(h, l);
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

__device__ __forceinline__ double slab_min(double x) {
  x = dmin(x, xor_get<1>(x));
  x = dmin(x, xor_get<2>(x));
  x = dmin(x, xor_get<4>(x));
  x = dmin(x, xor_get<8>(x));
  return x;
}
__device__ __forceinline__ double slab_max(double x) {
  x = dmax(x, xor_get<1>(x));
  x = dmax(x, xor_get<2>(x));
  x = dmax(x, xor_get<4>(x));
  x = dmax(x, xor_get<8>(x));
  return x;
}

// true if any lane of the caller's slab has pred set
__device__ __forceinline__ bool slab_any(bool pred) {
  const uint64_t b = __ballot(pred);
  const int shift = (__lane_id() & 48);
  return ((b >> shift) & 0xFFFFull) != 0;
}

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
