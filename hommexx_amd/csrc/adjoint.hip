// Adjoint (reverse-mode derivative) of the dry preqx dynamics step with
// rsplit = 0.  With rsplit = 0 the vertical coordinate is Eulerian, the
// vertical remap moves only tracers, and the step maps (v, T, ps) at n0 to
// (v, T, ps) at np1 through
//     dp3d = dhyai*ps0 + dhybi*ps                      (prim_driver.cpp:101-110)
//     five CAAR stages + DSS, RK5 combination          (prim_advance_exp.cpp:113-161)
//     hypervis_subcycle biharmonic cycles              (HyperviscosityFunctorImpl.hpp)
//     ps = hyai(0)*ps0 + sum_k dp3d                    (rsplit = 0 branch of the remap)
// Tracers do not act on the dynamics of a dry run, so they have no adjoint
// here.  The kernels below are the transposes of the forward kernels' linear
// maps, evaluated at the forward states the driver recomputes:
//   * element operators: the transpose of a 4-point contraction along a GLL
//     row (column) is the contraction with the transposed derivative matrix
//     along the same row (column), so the lane primitives of lanes.hpp serve;
//   * DSS with rspheremp is self-adjoint (rspheremp is the same at every copy
//     of a node and the copy-sum is symmetric), so the forward DSS kernel is
//     its own adjoint;
//   * column scans become reverse scans (prefix <-> suffix).
// Operation order is free here: gradients are checked against a tolerance,
// not bit for bit.
#include "kernels.hpp"

namespace hx {

namespace {

// ------------------------------------------------- transposed element operators
// adjoint of gradient_sphere: (g0b, g1b) -> sb
__device__ __forceinline__ double grad_T(const ElemLane &e, const DvvLane &d, double g0b, double g1b) {
  const double v0b = e.Dinv[0] * g0b + e.Dinv[2] * g1b;
  const double v1b = e.Dinv[1] * g0b + e.Dinv[3] * g1b;
  return (row_contract(d.TJ, v0b) + col_contract(d.TI, v1b)) * RREARTH;
}

// adjoint of divergence_sphere: db -> (ab, bb)
__device__ __forceinline__ void div_T(const ElemLane &e, const DvvLane &d, double db, double &ab, double &bb) {
  const double w = db * (1.0 / e.metdet * RREARTH);
  const double g0 = row_contract(d.TJ, w);
  const double g1 = col_contract(d.TI, w);
  ab = (e.Dinv[0] * g0 + e.Dinv[1] * g1) * e.metdet;
  bb = (e.Dinv[2] * g0 + e.Dinv[3] * g1) * e.metdet;
}

// adjoint of vorticity_sphere: zb -> (ub, vb)
__device__ __forceinline__ void vort_T(const ElemLane &e, const DvvLane &d, double zb, double &ub, double &vb) {
  const double w = zb * (1.0 / e.metdet * RREARTH);
  const double c1 = row_contract(d.TJ, w);
  const double c0 = -col_contract(d.TI, w);
  ub = e.D[0] * c0 + e.D[2] * c1;
  vb = e.D[1] * c0 + e.D[3] * c1;
}

// adjoint of divergence_sphere_wk: ddb -> (v0b, v1b)
__device__ __forceinline__ void div_wk_T(const ElemLane &e, const DvvLane &d, double ddb, double &v0b,
                                         double &v1b) {
  const double ab = -RREARTH * e.sph * row_contract(d.J, ddb);
  const double bb = -RREARTH * e.sph * col_contract(d.I, ddb);
  v0b = e.Dinv[0] * ab + e.Dinv[1] * bb;
  v1b = e.Dinv[2] * ab + e.Dinv[3] * bb;
}

// adjoint of grad_sphere_wk_testcov: (g0b, g1b) -> sb
__device__ __forceinline__ double gwk_T(const ElemLane &e, const DvvLane &d, double g0b, double g1b) {
  const double b0 = (e.D[0] * g0b + e.D[1] * g1b) * RREARTH;
  const double b1 = (e.D[2] * g0b + e.D[3] * g1b) * RREARTH;
  const double P = e.metdet * (e.metinv[0] * b0 + e.metinv[2] * b1);
  const double Q = e.metdet * (e.metinv[1] * b0 + e.metinv[3] * b1);
  return -e.mp * (row_contract(d.J, P) + col_contract(d.I, Q));
}

// adjoint of curl_sphere_wk_testcov: (c0b, c1b) -> sb
__device__ __forceinline__ double cwk_T(const ElemLane &e, const DvvLane &d, double c0b, double c1b) {
  const double s0 = (e.D[0] * c0b + e.D[1] * c1b) * RREARTH;
  const double s1 = (e.D[2] * c0b + e.D[3] * c1b) * RREARTH;
  return e.mp * (row_contract(d.J, s1) - col_contract(d.I, s0));
}

// adjoint of laplace_simple
__device__ __forceinline__ double lap_T(const ElemLane &e, const DvvLane &d, double lb) {
  double g0b, g1b;
  div_wk_T(e, d, lb, g0b, g1b);
  return grad_T(e, d, g0b, g1b);
}

// adjoint of vlaplace_sphere_wk_contra
__device__ __forceinline__ void vlap_T(const ElemLane &e, const DvvLane &d, double nu_ratio, double l0b, double l1b,
                                       double &u0b, double &u1b) {
  const double f = 2.0 * e.sph * RREARTH * RREARTH;
  double db = gwk_T(e, d, l0b, l1b);
  if (nu_ratio > 0 && nu_ratio != 1.0) db *= nu_ratio;
  double a, b, c, w;
  div_T(e, d, db, a, b);
  vort_T(e, d, -cwk_T(e, d, l0b, l1b), c, w);
  u0b = f * l0b + a + c;
  u1b = f * l1b + b + w;
}

// ------------------------------------------------------------ lane mapping
// slab kernels: 16 lanes per (element, level), levels fastest
__device__ __forceinline__ bool adj_slab(const AdjCaarArgs &a, int &ie, int &k, int &pt) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s = t >> 4;
  pt = (int)(t & 15);
  ie = (int)(s / a.nlev);
  k = (int)(s % a.nlev);
  return ie < a.nelem;
}
__device__ __forceinline__ size_t sidx(int nlev, int ie, int k, int pt) {
  return ((size_t)ie * nlev + k) * 16 + pt;
}
__device__ __forceinline__ size_t vidx(int nlev, int ie, int k, int c, int pt) {
  return (((size_t)ie * nlev + k) * 2 + c) * 16 + pt;
}

// ------------------------------------------------------------- CAAR: forward
// F1: div(v dp) per slab
__global__ __launch_bounds__(256) void adj_caar_div_kernel(AdjCaarArgs a) {
  int ie, k, pt;
  if (!adj_slab(a, ie, k, pt)) return;
  ElemLane e;
  DvvLane d;
  load_elem(a.m, ie, pt, e);
  load_dvv(a.m.dvv, pt, d);
  const int L = a.nlev;
  const double dp = a.dp[sidx(L, ie, k, pt)];
  const double u0 = a.v[vidx(L, ie, k, 0, pt)], u1 = a.v[vidx(L, ie, k, 1, pt)];
  a.div[sidx(L, ie, k, pt)] = sph_div(e, d, u0 * dp, u1 * dp);
}

// F2: pressure p, exclusive prefix C of div(v dp), column total
__global__ __launch_bounds__(256) void adj_caar_scan_kernel(AdjCaarArgs a) {
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (size_t)a.nelem * 16) return;
  const int ie = (int)(col >> 4), pt = (int)(col & 15), L = a.nlev;
  double p = a.p_top, dprev = 0.0, c = 0.0;
  for (int k = 0; k < L; ++k) {
    const size_t i = sidx(L, ie, k, pt);
    const double dp = a.dp[i];
    p = p + 0.5 * (dprev + dp);
    dprev = dp;
    a.p[i] = p;
    a.c[i] = c;
    c += a.div[i];
  }
  a.stot[col] = c;
}

// ------------------------------------------------------------ CAAR: adjoint
// B1: pointwise and horizontal transposes at one level, vertical-advection
// adjoint in gather form (reads levels k-1 and k+1)
__global__ __launch_bounds__(256) void adj_caar_b1_kernel(AdjCaarArgs a) {
  int ie, k, pt;
  if (!adj_slab(a, ie, k, pt)) return;
  ElemLane e;
  DvvLane d;
  load_elem(a.m, ie, pt, e);
  load_dvv(a.m.dvv, pt, d);
  const int L = a.nlev;
  const size_t i = sidx(L, ie, k, pt);
  const double h = a.h, sph = e.sph;
  const double u0 = a.v[vidx(L, ie, k, 0, pt)], u1 = a.v[vidx(L, ie, k, 1, pt)];
  const double T = a.t[i], dp = a.dp[i], P = a.p[i];
  const double aT = h * sph * a.lp_t[i];
  const double au0 = h * sph * a.lp_v[vidx(L, ie, k, 0, pt)];
  const double au1 = h * sph * a.lp_v[vidx(L, ie, k, 1, pt)];
  const double adp = h * sph * a.lp_dp[i];
  double gp0, gp1, gT0, gT1;
  sph_grad(e, d, P, gp0, gp1);
  sph_grad(e, d, T, gT0, gT1);
  const double vort = sph_vort(e, d, u0, u1) + e.fcor;
  const double div = a.div[i], C = a.c[i];
  const double omb = (u0 * gp0 + u1 * gp1 - (C + 0.5 * div)) / P;
  const double lomb = KAPPA * T * aT;
  const double leg0 = -au0, leg1 = -au1;
  const double lvort = au0 * u1 - au1 * u0;
  const double lgp0 = RGAS * (T / P) * leg0 + u0 * lomb / P;
  const double lgp1 = RGAS * (T / P) * leg1 + u1 * lomb / P;
  const double gpl = gp0 * leg0 + gp1 * leg1;
  double lT = KAPPA * omb * aT + RGAS / P * gpl + grad_T(e, d, -aT * u0, -aT * u1);
  const double lephi = grad_T(e, d, leg0, leg1);
  double hv0, hv1;
  vort_T(e, d, lvort, hv0, hv1);
  double lu0 = -aT * gT0 - au1 * vort + u0 * lephi + gp0 * lomb / P + hv0;
  double lu1 = -aT * gT1 + au0 * vort + u1 * lephi + gp1 * lomb / P + hv1;
  a.pb[i] = -RGAS * T / (P * P) * gpl - omb / P * lomb + grad_T(e, d, lgp0, lgp1);
  a.phib[i] = lephi;
  a.ombp[i] = lomb / P;
  // vertical advection: ta(j) = facp(j)*(f(j+1)-f(j)) + facm(j)*(f(j)-f(j-1)),
  // facp(j) = 0.5/dp(j)*eta(j+1), facm(j) = 0.5/dp(j)*eta(j), eta(0) = eta(L) = 0
  const double stot = a.stot[(size_t)ie * 16 + pt];
  const double eta_k = k == 0 ? 0.0 : a.hybi[k] * stot - C;
  const double eta_k1 = k + 1 >= L ? 0.0 : a.hybi[k + 1] * stot - a.c[sidx(L, ie, k + 1, pt)];
  const double facp = 0.5 / dp * eta_k1, facm = 0.5 / dp * eta_k;
  // adjoint of the T and v vertical-advection terms: la = -(h sph lambda_pre)
  const double laT = -aT, la0 = -au0, la1 = -au1;
  double lfacp = 0.0, lfacm = 0.0, lfacp_m = 0.0;
  double oT = (facm - facp) * laT, o0 = (facm - facp) * la0, o1 = (facm - facp) * la1;
  double dp_m = 1.0, adp_m = 0.0;
  if (k + 1 < L) {
    const size_t j = sidx(L, ie, k + 1, pt);
    const double dpp = a.dp[j];
    const double facm_p = 0.5 / dpp * eta_k1;
    const double hs = h * sph;
    const double lT_p = -hs * a.lp_t[j], l0_p = -hs * a.lp_v[vidx(L, ie, k + 1, 0, pt)],
                 l1_p = -hs * a.lp_v[vidx(L, ie, k + 1, 1, pt)];
    oT -= facm_p * lT_p;
    o0 -= facm_p * l0_p;
    o1 -= facm_p * l1_p;
    lfacp = (a.t[j] - T) * laT + (a.v[vidx(L, ie, k + 1, 0, pt)] - u0) * la0 +
            (a.v[vidx(L, ie, k + 1, 1, pt)] - u1) * la1;
  }
  if (k > 0) {
    const size_t j = sidx(L, ie, k - 1, pt);
    dp_m = a.dp[j];
    const double facp_m = 0.5 / dp_m * eta_k;
    const double hs = h * sph;
    const double lT_m = -hs * a.lp_t[j], l0_m = -hs * a.lp_v[vidx(L, ie, k - 1, 0, pt)],
                 l1_m = -hs * a.lp_v[vidx(L, ie, k - 1, 1, pt)];
    adp_m = hs * a.lp_dp[j];
    oT += facp_m * lT_m;
    o0 += facp_m * l0_m;
    o1 += facp_m * l1_m;
    const double dT = T - a.t[j], d0 = u0 - a.v[vidx(L, ie, k - 1, 0, pt)], d1 = u1 - a.v[vidx(L, ie, k - 1, 1, pt)];
    lfacm = dT * laT + d0 * la0 + d1 * la1;
    lfacp_m = dT * lT_m + d0 * l0_m + d1 * l1_m;
  }
  lT += oT;
  lu0 += o0;
  lu1 += o1;
  a.lt[i] = lT;
  a.lv[vidx(L, ie, k, 0, pt)] = lu0;
  a.lv[vidx(L, ie, k, 1, pt)] = lu1;
  a.ldp[i] = -0.5 * eta_k1 / (dp * dp) * lfacp - 0.5 * eta_k / (dp * dp) * lfacm;
  a.etab[i] = k == 0 ? 0.0 : 0.5 / dp_m * lfacp_m + 0.5 / dp * lfacm + adp - adp_m;
  a.divb[i] = -adp;
}

// B2: column reverse scans: geopotential, pressure, eta and omega sums
__global__ __launch_bounds__(256) void adj_caar_b2_kernel(AdjCaarArgs a) {
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (size_t)a.nelem * 16) return;
  const int ie = (int)(col >> 4), pt = (int)(col & 15), L = a.nlev;
  double phisum = 0.0, lstot = 0.0;
  for (int k = 0; k < L; ++k) {
    const size_t i = sidx(L, ie, k, pt);
    const double phib = a.phib[i];
    const double lr = phib + 2.0 * phisum;
    phisum += phib;
    const double P = a.p[i], T = a.t[i], dp = a.dp[i];
    a.lt[i] += RGAS * dp * 0.5 / P * lr;
    a.ldp[i] += RGAS * T * 0.5 / P * lr;
    a.pb[i] -= RGAS * T * dp * 0.5 / (P * P) * lr;
    lstot += a.hybi[k] * a.etab[i];
  }
  double spb = 0.0, scb = 0.0;
  for (int k = L - 1; k >= 0; --k) {
    const size_t i = sidx(L, ie, k, pt);
    const double pb = a.pb[i];
    a.ldp[i] += 0.5 * pb + spb;
    spb += pb;
    const double ombp = a.ombp[i];
    a.divb[i] += scb + lstot - 0.5 * ombp;
    scb += -a.etab[i] - ombp;
  }
}

// B3: adjoint of div(v dp); results into the stage's adjoint accumulators
__global__ __launch_bounds__(256) void adj_caar_b3_kernel(AdjCaarArgs a) {
  int ie, k, pt;
  if (!adj_slab(a, ie, k, pt)) return;
  ElemLane e;
  DvvLane d;
  load_elem(a.m, ie, pt, e);
  load_dvv(a.m.dvv, pt, d);
  const int L = a.nlev;
  const size_t i = sidx(L, ie, k, pt);
  const size_t i0 = vidx(L, ie, k, 0, pt), i1 = vidx(L, ie, k, 1, pt);
  double ab, bb;
  div_T(e, d, a.divb[i], ab, bb);
  const double dp = a.dp[i], u0 = a.v[i0], u1 = a.v[i1];
  double lb[4] = {a.lv[i0] + dp * ab, a.lv[i1] + dp * bb, a.lt[i], a.ldp[i] + u0 * ab + u1 * bb};
  const double sph = e.sph;
  const double la[4] = {sph * a.lp_v[i0], sph * a.lp_v[i1], sph * a.lp_t[i], sph * a.lp_dp[i]};
  double *bp[4] = {a.b_v + i0, a.b_v + i1, a.b_t + i, a.b_dp + i};
  const bool alias = a.a1_v == a.b_v;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    double x = lb[f];
    if (a.b_acc) x += *bp[f];
    if (alias) x += a.c1 * la[f];
    *bp[f] = x;
  }
  if (!alias) {
    a.a1_v[i0] += a.c1 * la[0];
    a.a1_v[i1] += a.c1 * la[1];
    a.a1_t[i] += a.c1 * la[2];
    a.a1_dp[i] += a.c1 * la[3];
  }
  if (a.a2_v) {
    a.a2_v[i0] += a.c2 * la[0];
    a.a2_v[i1] += a.c2 * la[1];
    a.a2_t[i] += a.c2 * la[2];
    a.a2_dp[i] += a.c2 * la[3];
  }
}

// ------------------------------------------------------------- HV: adjoint
// heating and update: T'' = T + tt - (vt . v'')/cp, v'' = v + vt, dp'' = r*S(...)
__global__ __launch_bounds__(256) void adj_hv_pre_kernel(AdjHvArgs a) {
  const size_t n = (size_t)a.nelem * a.nlev * 16;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t ek = i >> 4, pt = i & 15;
  const size_t i0 = ek * 32 + pt, i1 = i0 + 16;
  const double lT = a.l_t[i];
  const double r = lT / CP;
  const double vo0 = a.xo_v[i0], vo1 = a.xo_v[i1];
  const double vt0 = vo0 - a.x_v[i0], vt1 = vo1 - a.x_v[i1];
  const double lt0 = a.l_v[i0] - r * vt0, lt1 = a.l_v[i1] - r * vt1;
  a.s_v[i0] = lt0 - r * vo0;
  a.s_v[i1] = lt1 - r * vo1;
  a.s_t[i] = lT;
  a.s_dp[i] = a.l_dp[i];
  a.l_v[i0] = lt0;
  a.l_v[i1] = lt1;
}

// after the first DSS: sponge terms, pointwise dp term and the transposed
// second laplacian (into the scratch for the second DSS)
__global__ __launch_bounds__(256) void adj_hv_a_kernel(AdjHvArgs a) {
  int ie, k, pt;
  {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long s = t >> 4;
    pt = (int)(t & 15);
    ie = (int)(s / a.nlev);
    k = (int)(s % a.nlev);
    if (ie >= a.nelem) return;
  }
  ElemLane e;
  DvvLane d;
  load_elem(a.m, ie, pt, e);
  load_dvv(a.m.dvv, pt, d);
  const int L = a.nlev;
  const size_t i = sidx(L, ie, k, pt), i0 = vidx(L, ie, k, 0, pt), i1 = vidx(L, ie, k, 1, pt);
  const double mv0 = a.dt * a.m_v[i0], mv1 = a.dt * a.m_v[i1], mT = a.dt * a.m_t[i], mdp = a.m_dp[i];
  const double mdpi = a.dt * mdp;
  double lv0 = a.l_v[i0], lv1 = a.l_v[i1], lT = a.l_t[i], ldp = e.sph * mdp;
  if (a.nu_top > 0 && k < 3) {
    const double sc = (k == 0 ? 4.0 : k == 1 ? 2.0 : 1.0) * a.nu_top;
    double x0, x1;
    vlap_T(e, d, 1.0, mv0, mv1, x0, x1);
    lv0 += sc * x0;
    lv1 += sc * x1;
    lT += sc * lap_T(e, d, mT);
    ldp += sc * lap_T(e, d, mdpi);
  }
  a.l_v[i0] = lv0;
  a.l_v[i1] = lv1;
  a.l_t[i] = lT;
  a.l_dp[i] = ldp;
  double z0, z1;
  vlap_T(e, d, a.nu_ratio2, -a.nu * mv0, -a.nu * mv1, z0, z1);
  a.s_v[i0] = z0;
  a.s_v[i1] = z1;
  a.s_t[i] = lap_T(e, d, -a.nu_s * mT);
  a.s_dp[i] = lap_T(e, d, -a.nu_p * mdpi);
}

// after the second DSS: the transposed first laplacian
__global__ __launch_bounds__(256) void adj_hv_b_kernel(AdjHvArgs a) {
  int ie, k, pt;
  {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long s = t >> 4;
    pt = (int)(t & 15);
    ie = (int)(s / a.nlev);
    k = (int)(s % a.nlev);
    if (ie >= a.nelem) return;
  }
  ElemLane e;
  DvvLane d;
  load_elem(a.m, ie, pt, e);
  load_dvv(a.m.dvv, pt, d);
  const int L = a.nlev;
  const size_t i = sidx(L, ie, k, pt), i0 = vidx(L, ie, k, 0, pt), i1 = vidx(L, ie, k, 1, pt);
  double x0, x1;
  vlap_T(e, d, a.nu_ratio1, a.m_v[i0], a.m_v[i1], x0, x1);
  a.l_v[i0] += x0;
  a.l_v[i1] += x1;
  a.l_t[i] += lap_T(e, d, a.m_t[i]);
  a.l_dp[i] += lap_T(e, d, a.m_dp[i]);
}

// ------------------------------------------------------------- small kernels
__global__ void adj_fill_dp_kernel(int nelem, int nlev, const double *__restrict__ lps, double *__restrict__ ldp) {
  const size_t n = (size_t)nelem * nlev * 16;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t ie = i / ((size_t)nlev * 16);
  ldp[i] = lps[ie * 16 + (i & 15)];
}

__global__ void adj_ps_kernel(int nelem, int nlev, const double *__restrict__ hybi, const double *__restrict__ ldp,
                              double *__restrict__ lps) {
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (size_t)nelem * 16) return;
  const size_t ie = col >> 4, pt = col & 15;
  double s = 0.0;
  for (int k = 0; k < nlev; ++k) s += (hybi[k + 1] - hybi[k]) * ldp[(ie * nlev + k) * 16 + pt];
  lps[col] = s;
}

__global__ void adj_zero_kernel(size_t n, double *__restrict__ p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

inline unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_adj_caar(hipStream_t st, const AdjCaarArgs &a) {
  const size_t slab_threads = (size_t)a.nelem * a.nlev * 16;
  const size_t cols = (size_t)a.nelem * 16;
  hipLaunchKernelGGL(adj_caar_div_kernel, dim3(nblk(slab_threads)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_scan_kernel, dim3(nblk(cols)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_b1_kernel, dim3(nblk(slab_threads)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_b2_kernel, dim3(nblk(cols)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_b3_kernel, dim3(nblk(slab_threads)), dim3(256), 0, st, a);
}

void launch_adj_hv_pre(hipStream_t st, const AdjHvArgs &a) {
  hipLaunchKernelGGL(adj_hv_pre_kernel, dim3(nblk((size_t)a.nelem * a.nlev * 16)), dim3(256), 0, st, a);
}
void launch_adj_hv_a(hipStream_t st, const AdjHvArgs &a) {
  hipLaunchKernelGGL(adj_hv_a_kernel, dim3(nblk((size_t)a.nelem * a.nlev * 16)), dim3(256), 0, st, a);
}
void launch_adj_hv_b(hipStream_t st, const AdjHvArgs &a) {
  hipLaunchKernelGGL(adj_hv_b_kernel, dim3(nblk((size_t)a.nelem * a.nlev * 16)), dim3(256), 0, st, a);
}
void launch_adj_fill_dp(hipStream_t st, int nelem, int nlev, const double *lps, double *ldp) {
  hipLaunchKernelGGL(adj_fill_dp_kernel, dim3(nblk((size_t)nelem * nlev * 16)), dim3(256), 0, st, nelem, nlev, lps,
                     ldp);
}
void launch_adj_ps(hipStream_t st, int nelem, int nlev, const double *hybi, const double *ldp, double *lps) {
  hipLaunchKernelGGL(adj_ps_kernel, dim3(nblk((size_t)nelem * 16)), dim3(256), 0, st, nelem, nlev, hybi, ldp, lps);
}
void launch_adj_zero(hipStream_t st, size_t n, double *p) {
  hipLaunchKernelGGL(adj_zero_kernel, dim3(nblk(n)), dim3(256), 0, st, n, p);
}

}  // namespace hx
