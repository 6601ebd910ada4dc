// FP32 spherical operators of the mixed form (hyperviscosity, tracer flux
// divergence, tracer biharmonic).  Same operator definitions and term order
// as sphere.hpp (SphereOperators.hpp in the reference), evaluated in single
// precision: tensor-product contractions use the compensated dot products of
// lanes.hpp, and multi-term weak-form sums carry a TwoSum error term.
#pragma once
#include "sphere.hpp"

namespace hx {

struct ElemLaneF {
  float Dinv[4], D[4], metinv[4];
  float metdet, rmetdet, sph;
  float mp_row[4], mp_col[4];
};

__device__ __forceinline__ void load_elemf(const MetricPtrs &m, int ie, int pt, ElemLaneF &e) {
  const size_t t = (size_t)ie * 64 + pt, s = (size_t)ie * 16 + pt;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    e.Dinv[k] = (float)m.Dinv[t + 16 * k];
    e.D[k] = (float)m.D[t + 16 * k];
    e.metinv[k] = (float)m.metinv[t + 16 * k];
  }
  const double md = m.metdet[s];
  e.metdet = (float)md;
  e.rmetdet = (float)(1.0 / md * RREARTH);
  e.sph = (float)m.spheremp[s];
  const int ig = pt >> 2, jg = pt & 3;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    e.mp_row[k] = (float)m.mp[(size_t)ie * 16 + ig * 4 + k];
    e.mp_col[k] = (float)m.mp[(size_t)ie * 16 + k * 4 + jg];
  }
}

constexpr float RREARTHF = (float)RREARTH;

// s += p with the rounding error of the sum kept in e (TwoSum)
__device__ __forceinline__ void acc2(float p, float &s, float &e) {
  const float t = s + p;
  const float z = t - s;
  e += (s - (t - z)) + (p - z);
  s = t;
}

__device__ __forceinline__ void sph_gradf(const ElemLaneF &e, const DvvLaneF &d, float s, float &g0, float &g1) {
  const float v0 = row_contractf(d.J, s) * RREARTHF;
  const float v1 = col_contractf(d.I, s) * RREARTHF;
  g0 = e.Dinv[0] * v0 + e.Dinv[1] * v1;
  g1 = e.Dinv[2] * v0 + e.Dinv[3] * v1;
}

__device__ __forceinline__ float sph_divf(const ElemLaneF &e, const DvvLaneF &d, float v0, float v1) {
  const float gv0 = (e.Dinv[0] * v0 + e.Dinv[2] * v1) * e.metdet;
  const float gv1 = (e.Dinv[1] * v0 + e.Dinv[3] * v1) * e.metdet;
  return (row_contractf(d.J, gv0) + col_contractf(d.I, gv1)) * e.rmetdet;
}

__device__ __forceinline__ float sph_vortf(const ElemLaneF &e, const DvvLaneF &d, float u, float v) {
  const float c0 = e.D[0] * u + e.D[1] * v;
  const float c1 = e.D[2] * u + e.D[3] * v;
  return (row_contractf(d.J, c1) - col_contractf(d.I, c0)) * e.rmetdet;
}

template <int K>
__device__ __forceinline__ void divwk_term(const DvvLaneF &d, float sa, float sb, float &s, float &err) {
  acc2(-(row_getf<K>(sa) * d.TJ[K]), s, err);
  acc2(-(col_getf<K>(sb) * d.TI[K]), s, err);
}
__device__ __forceinline__ float sph_div_wkf(const ElemLaneF &e, const DvvLaneF &d, float v0, float v1) {
  const float a = e.Dinv[0] * v0 + e.Dinv[2] * v1;
  const float b = e.Dinv[1] * v0 + e.Dinv[3] * v1;
  const float sa = e.sph * a, sb = e.sph * b;
  float s = 0.f, err = 0.f;
  divwk_term<0>(d, sa, sb, s, err);
  divwk_term<1>(d, sa, sb, s, err);
  divwk_term<2>(d, sa, sb, s, err);
  divwk_term<3>(d, sa, sb, s, err);
  return (s + err) * RREARTHF;
}

__device__ __forceinline__ float sph_laplacef(const ElemLaneF &e, const DvvLaneF &d, float f) {
  float g0, g1;
  sph_gradf(e, d, f, g0, g1);
  return sph_div_wkf(e, d, g0, g1);
}

template <int J>
__device__ __forceinline__ void gwk_termf(const ElemLaneF &e, const DvvLaneF &d, float s, float &b0, float &e0,
                                          float &b1, float &e1) {
  const float md = e.metdet;
  const float snj = row_getf<J>(s) * d.TJ[J] * e.mp_row[J] * md;
  const float sjm = col_getf<J>(s) * d.TI[J] * e.mp_col[J] * md;
  acc2(-(e.metinv[0] * snj), b0, e0);
  acc2(-(e.metinv[1] * sjm), b0, e0);
  acc2(-(e.metinv[2] * snj), b1, e1);
  acc2(-(e.metinv[3] * sjm), b1, e1);
}
__device__ __forceinline__ void sph_grad_wk_testcovf(const ElemLaneF &e, const DvvLaneF &d, float s, float &g0,
                                                     float &g1) {
  float b0 = 0.f, e0 = 0.f, b1 = 0.f, e1 = 0.f;
  gwk_termf<0>(e, d, s, b0, e0, b1, e1);
  gwk_termf<1>(e, d, s, b0, e0, b1, e1);
  gwk_termf<2>(e, d, s, b0, e0, b1, e1);
  gwk_termf<3>(e, d, s, b0, e0, b1, e1);
  b0 += e0;
  b1 += e1;
  g0 = (e.D[0] * b0 + e.D[2] * b1) * RREARTHF;
  g1 = (e.D[1] * b0 + e.D[3] * b1) * RREARTHF;
}

template <int J>
__device__ __forceinline__ void cwk_termf(const ElemLaneF &e, const DvvLaneF &d, float s, float &b0, float &e0,
                                          float &b1, float &e1) {
  acc2(-(e.mp_col[J] * col_getf<J>(s) * d.TI[J]), b0, e0);
  acc2(e.mp_row[J] * row_getf<J>(s) * d.TJ[J], b1, e1);
}
// c += -(curl_sphere_wk_testcov(s))  (alpha = -1, beta = 1)
__device__ __forceinline__ void sph_curl_wk_subf(const ElemLaneF &e, const DvvLaneF &d, float s, float &c0,
                                                 float &c1) {
  float b0 = 0.f, e0 = 0.f, b1 = 0.f, e1 = 0.f;
  cwk_termf<0>(e, d, s, b0, e0, b1, e1);
  cwk_termf<1>(e, d, s, b0, e0, b1, e1);
  cwk_termf<2>(e, d, s, b0, e0, b1, e1);
  cwk_termf<3>(e, d, s, b0, e0, b1, e1);
  b0 += e0;
  b1 += e1;
  c0 = c0 - (e.D[0] * b0 + e.D[2] * b1) * RREARTHF;
  c1 = c1 - (e.D[1] * b0 + e.D[3] * b1) * RREARTHF;
}

// vlaplace_sphere_wk_contra
__device__ __forceinline__ void sph_vlaplace_contraf(const ElemLaneF &e, const DvvLaneF &d, float nu_ratio, float u0,
                                                     float u1, float &l0, float &l1) {
  float div = sph_divf(e, d, u0, u1);
  if (nu_ratio > 0.f && nu_ratio != 1.f) div *= nu_ratio;
  float gc0, gc1;
  sph_grad_wk_testcovf(e, d, div, gc0, gc1);
  const float vort = sph_vortf(e, d, u0, u1);
  sph_curl_wk_subf(e, d, vort, gc0, gc1);
  const float re2 = RREARTHF * RREARTHF;
  const float f = 2.f * e.sph;
  l0 = f * u0 * re2 + gc0;
  l1 = f * u1 * re2 + gc1;
}

// scalar value of lane 0 of the caller's slab (a per-slab reference value:
// laplacians annihilate constants, so subtracting it before rounding to FP32
// keeps the relative precision of the variation)
__device__ __forceinline__ double slab_ref(double x) {
  constexpr int pattern = 0x10 | (0x00 << 5);  // and 0x10: lane 0 of the 16-lane row
  const int l = __builtin_amdgcn_ds_swizzle(lo32(x), pattern);
  const int h = __builtin_amdgcn_ds_swizzle(hi32(x), pattern);
  return mk64(h, l);
}

}  // namespace hx
