// np=4 spherical operators evaluated by a 16-lane slab (one element, one
// level).  Each function returns the caller lane's value.  The arithmetic
// follows SphereOperators.hpp (reference) term by term: products left to
// right, contractions summed k = 0..3 from zero, so results match a serial
// CPU evaluation bit for bit when compiled without FMA contraction.
#pragma once
#include "lanes.hpp"

namespace hx {

// Metric at the lane's point.  Tensors use the reference C++ order
// T(idim, jdim) = t[idim*2 + jdim], which is the flat Fortran array.
struct ElemLane {
  double Dinv[4], D[4], metinv[4];
  double metdet, sph, rsmp, mp, fcor, phis;
  double mp_row[4];  // mp(igp, k)
  double mp_col[4];  // mp(k, jgp)
};

struct MetricPtrs {
  const double *D, *Dinv, *metinv;           // [ie][4][16]
  const double *fcor, *mp, *spheremp, *rspheremp, *metdet, *phis;  // [ie][16]
  const double *dvv;                          // [4][4], dvv[a][b] = L_b'(x_a)
  const double *tensorvisc, *vec_sph2cart;    // [ie][4][16], [ie][6][16] (tensor hyperviscosity)
};

__device__ __forceinline__ void load_elem(const MetricPtrs &m, int ie, int pt, ElemLane &e) {
  const size_t t = (size_t)ie * 64 + pt, s = (size_t)ie * 16 + pt;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    e.Dinv[k] = m.Dinv[t + 16 * k];
    e.D[k] = m.D[t + 16 * k];
    e.metinv[k] = m.metinv[t + 16 * k];
  }
  e.metdet = m.metdet[s];
  e.sph = m.spheremp[s];
  e.rsmp = m.rspheremp[s];
  e.mp = m.mp[s];
  e.fcor = m.fcor[s];
  e.phis = m.phis[s];
  const int ig = pt >> 2, jg = pt & 3;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    e.mp_row[k] = m.mp[(size_t)ie * 16 + ig * 4 + k];
    e.mp_col[k] = m.mp[(size_t)ie * 16 + k * 4 + jg];
  }
}

// gradient_sphere (SphereOperators.hpp:291-319)
__device__ __forceinline__ void sph_grad(const ElemLane &e, const DvvLane &d, double s, double &g0, double &g1) {
  double v0 = row_contract(d.J, s);
  double v1 = col_contract(d.I, s);
  v0 *= RREARTH;
  v1 *= RREARTH;
  g0 = e.Dinv[0] * v0 + e.Dinv[1] * v1;
  g1 = e.Dinv[2] * v0 + e.Dinv[3] * v1;
}

// gradient_sphere_update (:321-348): g += grad(s)
__device__ __forceinline__ void sph_grad_update(const ElemLane &e, const DvvLane &d, double s, double &g0,
                                                double &g1) {
  double dsdx = row_contract(d.J, s);
  double dsdy = col_contract(d.I, s);
  dsdx *= RREARTH;
  dsdy *= RREARTH;
  g0 += e.Dinv[0] * dsdx + e.Dinv[1] * dsdy;
  g1 += e.Dinv[2] * dsdx + e.Dinv[3] * dsdy;
}

// divergence_sphere (:350-392)
__device__ __forceinline__ double sph_div(const ElemLane &e, const DvvLane &d, double v0, double v1) {
  const double gv0 = (e.Dinv[0] * v0 + e.Dinv[2] * v1) * e.metdet;
  const double gv1 = (e.Dinv[1] * v0 + e.Dinv[3] * v1) * e.metdet;
  const double dudx = row_contract(d.J, gv0);
  const double dvdy = col_contract(d.I, gv1);
  return (dudx + dvdy) * (1.0 / e.metdet * RREARTH);
}

// divergence_sphere_update (:396-444): qdp + alpha*div(vstar*qdp) + q0
__device__ __forceinline__ double sph_div_update(const ElemLane &e, const DvvLane &d, double alpha, double vs0,
                                                 double vs1, double qdp, double q0) {
  const double a = vs0 * qdp, b = vs1 * qdp;
  const double gv0 = (e.Dinv[0] * a + e.Dinv[2] * b) * e.metdet;
  const double gv1 = (e.Dinv[1] * a + e.Dinv[3] * b) * e.metdet;
  const double dudx = row_contract(d.J, gv0);
  const double dvdy = col_contract(d.I, gv1);
  return (qdp + alpha * ((dudx + dvdy) * (1.0 / e.metdet * RREARTH)) + q0);
}

// vorticity_sphere (:446-533)
__device__ __forceinline__ double sph_vort(const ElemLane &e, const DvvLane &d, double u, double v) {
  const double c0 = e.D[0] * u + e.D[1] * v;
  const double c1 = e.D[2] * u + e.D[3] * v;
  const double dvdx = row_contract(d.J, c1);
  const double dudy = col_contract(d.I, c0);
  return (dvdx - dudy) * (1.0 / e.metdet * RREARTH);
}

// divergence_sphere_wk (:536-583)
__device__ __forceinline__ double sph_div_wk(const ElemLane &e, const DvvLane &d, double v0, double v1) {
  const double a = e.Dinv[0] * v0 + e.Dinv[2] * v1;
  const double b = e.Dinv[1] * v0 + e.Dinv[3] * v1;
  const double sa = e.sph * a, sb = e.sph * b;
  double dd = 0;
  dd -= (row_get<0>(sa) * d.TJ[0] + col_get<0>(sb) * d.TI[0]) * RREARTH;
  dd -= (row_get<1>(sa) * d.TJ[1] + col_get<1>(sb) * d.TI[1]) * RREARTH;
  dd -= (row_get<2>(sa) * d.TJ[2] + col_get<2>(sb) * d.TI[2]) * RREARTH;
  dd -= (row_get<3>(sa) * d.TJ[3] + col_get<3>(sb) * d.TI[3]) * RREARTH;
  return dd;
}

// laplace_simple (:586-597)
__device__ __forceinline__ double sph_laplace(const ElemLane &e, const DvvLane &d, double f) {
  double g0, g1;
  sph_grad(e, d, f, g0, g1);
  return sph_div_wk(e, d, g0, g1);
}

// grad_sphere_wk_testcov (:712-748)
template <int J>
__device__ __forceinline__ void gwk_term(const ElemLane &e, const DvvLane &d, double s, double &b0, double &b1) {
  const double md = e.metdet;
  const double snj = row_get<J>(s), sjm = col_get<J>(s);
  const double mpnj = e.mp_row[J], mpjm = e.mp_col[J];
  const double djm = d.TJ[J], djn = d.TI[J];
  b0 -= (mpnj * e.metinv[0] * md * snj * djm + mpjm * e.metinv[1] * md * sjm * djn);
  b1 -= (mpnj * e.metinv[2] * md * snj * djm + mpjm * e.metinv[3] * md * sjm * djn);
}
__device__ __forceinline__ void sph_grad_wk_testcov(const ElemLane &e, const DvvLane &d, double s, double &g0,
                                                    double &g1) {
  double b0 = 0, b1 = 0;
  gwk_term<0>(e, d, s, b0, b1);
  gwk_term<1>(e, d, s, b0, b1);
  gwk_term<2>(e, d, s, b0, b1);
  gwk_term<3>(e, d, s, b0, b1);
  g0 = (e.D[0] * b0 + e.D[2] * b1) * RREARTH;
  g1 = (e.D[1] * b0 + e.D[3] * b1) * RREARTH;
}

// curl_sphere_wk_testcov_update (:671-710) with alpha = -1, beta = 1
template <int J>
__device__ __forceinline__ void cwk_term(const ElemLane &e, const DvvLane &d, double s, double &sb0, double &sb1) {
  sb0 -= e.mp_col[J] * col_get<J>(s) * d.TI[J];
  sb1 += e.mp_row[J] * row_get<J>(s) * d.TJ[J];
}
__device__ __forceinline__ void sph_curl_wk_update(const ElemLane &e, const DvvLane &d, double alpha, double beta,
                                                   double s, double &c0, double &c1) {
  double sb0 = 0, sb1 = 0;
  cwk_term<0>(e, d, s, sb0, sb1);
  cwk_term<1>(e, d, s, sb0, sb1);
  cwk_term<2>(e, d, s, sb0, sb1);
  cwk_term<3>(e, d, s, sb0, sb1);
  c0 = beta * c0 + alpha * (e.D[0] * sb0 + e.D[2] * sb1) * RREARTH;
  c1 = beta * c1 + alpha * (e.D[1] * sb0 + e.D[3] * sb1) * RREARTH;
}

// vlaplace_sphere_wk_contra (:816-862)
__device__ __forceinline__ void sph_vlaplace_contra(const ElemLane &e, const DvvLane &d, double nu_ratio, double u0,
                                                    double u1, double &l0, double &l1) {
  double div = sph_div(e, d, u0, u1);
  if (nu_ratio > 0 && nu_ratio != 1.0) div *= nu_ratio;
  double gc0, gc1;
  sph_grad_wk_testcov(e, d, div, gc0, gc1);
  const double vort = sph_vort(e, d, u0, u1);
  sph_curl_wk_update(e, d, -1.0, 1.0, vort, gc0, gc1);
  const double re2 = RREARTH * RREARTH;
  const double f = 2.0 * e.sph;
  l0 = f * u0 * re2 + gc0;
  l1 = f * u1 * re2 + gc1;
}

// tensor hyperviscosity: T(idim,jdim) = tv[idim*2+jdim], vec_sph2cart(c,i) = vs[c*3+i]
__device__ __forceinline__ void load_tensor(const MetricPtrs &m, int ie, int pt, double tv[4], double vs[6]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) tv[k] = m.tensorvisc[(size_t)ie * 64 + 16 * k + pt];
#pragma unroll
  for (int k = 0; k < 6; ++k) vs[k] = m.vec_sph2cart[(size_t)ie * 96 + 16 * k + pt];
}

// laplace_tensor (SphereOperators.hpp:602-635)
__device__ __forceinline__ double sph_laplace_tensor(const ElemLane &e, const DvvLane &d, const double tv[4], double f) {
  double g0, g1;
  sph_grad(e, d, f, g0, g1);
  const double s0 = tv[0] * g0 + tv[2] * g1;
  const double s1 = tv[1] * g0 + tv[3] * g1;
  return sph_div_wk(e, d, s0, s1);
}

// vlaplace_sphere_wk_cartesian (:750-814): tensor laplacians of the three
// cartesian components, projected back, plus the undamping 2 spheremp v / a^2
__device__ __forceinline__ void sph_vlaplace_cartesian(const ElemLane &e, const DvvLane &d, const double tv[4],
                                                       const double vs[6], double v0, double v1, double &l0,
                                                       double &l1) {
  const double a = sph_laplace_tensor(e, d, tv, vs[0] * v0 + vs[3] * v1);
  const double b = sph_laplace_tensor(e, d, tv, vs[1] * v0 + vs[4] * v1);
  const double c = sph_laplace_tensor(e, d, tv, vs[2] * v0 + vs[5] * v1);
  l0 = vs[0] * a + vs[1] * b + vs[2] * c + 2.0 * e.sph * v0 * (RREARTH) * (RREARTH);
  l1 = vs[3] * a + vs[4] * b + vs[5] * c + 2.0 * e.sph * v1 * (RREARTH) * (RREARTH);
}

}  // namespace hx
