// Adjoint (reverse-mode derivative) of the dry preqx dynamics step, rsplit
// 0 or 1.  The step maps (v, T, ps) at n0 to (v, T, ps) at np1 through
//     dp3d = dhyai*ps0 + dhybi*ps                      (prim_driver.cpp:101-110)
//     five CAAR stages + DSS, RK5 combination          (prim_advance_exp.cpp:113-161)
//     hypervis_subcycle biharmonic cycles              (HyperviscosityFunctorImpl.hpp)
//     rsplit = 0: ps = hyai(0)*ps0 + sum_k dp3d; the remap moves only tracers
//     rsplit = 1: PPM remap of v, T to the hybrid levels of that ps
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
  const size_t o = (size_t)ie * L * 16 + pt;
  const double *__restrict__ dpc = a.dp + o;
  const double *__restrict__ dvc = a.div + o;
  double *__restrict__ pc = a.p + o;
  double *__restrict__ cc = a.c + o;
  double p = a.p_top, dprev = 0.0, c = 0.0;
#pragma unroll 8
  for (int k = 0; k < L; ++k) {
    const double dp = dpc[(size_t)k * 16];
    p = p + 0.5 * (dprev + dp);
    dprev = dp;
    pc[(size_t)k * 16] = p;
    cc[(size_t)k * 16] = c;
    c += dvc[(size_t)k * 16];
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
  const bool rs0 = a.rsplit0;  // rsplit > 0: Lagrangian levels, no vertical flux
  const double eta_k = k == 0 || !rs0 ? 0.0 : a.hybi[k] * stot - C;
  const double eta_k1 = k + 1 >= L || !rs0 ? 0.0 : a.hybi[k + 1] * stot - a.c[sidx(L, ie, k + 1, pt)];
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
  a.etab[i] = k == 0 || !rs0 ? 0.0 : 0.5 / dp_m * lfacp_m + 0.5 / dp * lfacm + adp - adp_m;
  a.divb[i] = -adp;
}

// B2: column reverse scans: geopotential, pressure, eta and omega sums
// The column's levels are split over the 4 waves of a block (64 columns per
// block): each chunk re-runs the serial sums over the levels above it (phi
// terms, and the full column sum lstot) or below it (pb, the eta and omega
// terms), so every sum keeps its order and the result is that of one thread
// per column, with four times the threads in flight.
constexpr int B2C = 4;
__global__ __launch_bounds__(64 * B2C) void adj_caar_b2_kernel(AdjCaarArgs a) {
  const int chunk = threadIdx.x >> 6;
  const size_t col = (size_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const bool on = col < (size_t)a.nelem * 16;
  const int L = a.nlev, cs = (L + B2C - 1) / B2C;
  const int k0 = chunk * cs < L ? chunk * cs : L;
  const int k1 = k0 + cs < L ? k0 + cs : L;
  const size_t c = on ? col : 0;
  const int ie = (int)(c >> 4), pt = (int)(c & 15);
  const size_t o = (size_t)ie * L * 16 + pt;
  const double *__restrict__ phib = a.phib + o;
  const double *__restrict__ pc = a.p + o;
  const double *__restrict__ tc = a.t + o;
  const double *__restrict__ dpc = a.dp + o;
  const double *__restrict__ etab = a.etab + o;
  const double *__restrict__ ombp = a.ombp + o;
  double *__restrict__ lt = a.lt + o;
  double *__restrict__ ldp = a.ldp + o;
  double *__restrict__ pb = a.pb + o;
  double *__restrict__ divb = a.divb + o;
  double lstot = 0.0;
  if (on) {
    double phisum = 0.0;
#pragma unroll 8
    for (int k = 0; k < k0; ++k) phisum += phib[(size_t)k * 16];
#pragma unroll 8
    for (int k = 0; k < L; ++k) lstot += a.hybi[k] * etab[(size_t)k * 16];
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
      const size_t i = (size_t)k * 16;
      const double ph = phib[i];
      const double lr = ph + 2.0 * phisum;
      phisum += ph;
      const double P = pc[i], T = tc[i], dp = dpc[i];
      lt[i] += RGAS * dp * 0.5 / P * lr;
      ldp[i] += RGAS * T * 0.5 / P * lr;
      pb[i] -= RGAS * T * dp * 0.5 / (P * P) * lr;
    }
  }
  __syncthreads();  // pb of the levels below each chunk is final
  if (on) {
    double spb = 0.0, scb = 0.0;
#pragma unroll 8
    for (int k = L - 1; k >= k1; --k) {
      const size_t i = (size_t)k * 16;
      spb += pb[i];
      scb += -etab[i] - ombp[i];
    }
#pragma unroll 4
    for (int k = k1 - 1; k >= k0; --k) {
      const size_t i = (size_t)k * 16;
      const double b = pb[i];
      ldp[i] += 0.5 * b + spb;
      spb += b;
      const double om = ombp[i];
      divb[i] += scb + lstot - 0.5 * om;
      scb += -etab[i] - om;
    }
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

// ----------------------------------------------------------- remap adjoint
// Adjoint of the PPM remap of the states v (two components) and T for
// rsplit > 0 (PpmRemap.hpp:288-673; the forward is remap_kernel): extensive
// values f*dp are remapped from the source thicknesses dp (the state after
// hyperviscosity) to the hybrid levels of ps = hyai(0) ps0 + sum dp and
// divided by the target thicknesses.  The linearisation freezes every
// branch of the forward pass (the target search kid, the dma selection and
// the parabola adjustments): they are recomputed here from the same values.
// Scratch per column: REMAP_GRID arrays shared by the fields and
// REMAP_FIELD arrays per field, each nlev+4 long, stored [array][lev][col].
enum { RG_TG, RG_PIO, RG_PIN, RG_KID, RG_Z2 };
enum { RF_AA, RF_DMA, RF_AI, RF_LA, RF_LDMA, RF_LAI, RF_LC0, RF_LC1, RF_LC2, RF_LMASS, RF_LDX, RF_LD, RF_LPIO,
       RF_LPIN, RF_LTG };

struct RCol {
  const AdjRemapArgs *a;
  size_t col, ncol, W;
  __device__ double &g(int arr, int i) const { return a->scr[((size_t)arr * W + i) * ncol + col]; }
  __device__ double &f(int fld, int arr, int i) const {
    return a->scr[(size_t)REMAP_GRID * W * ncol + (((size_t)arr * 3 + fld) * W + i) * ncol + col];
  }
};

// grid factors of PpmRemap.hpp:367-413 at j from X = dx(j..j+3), with their
// partial derivatives dG[m] = dG/dX[m].  One division per factor: the value
// is N/Dn correctly rounded (div_r from the reciprocal r = 1/Dn, the same
// value as the forward remap's division); the derivatives are products with r
// (N/Dn^2 as (N/Dn)*r: within an ulp or two of the quotients, which the
// gradient's tolerance does not see).
__device__ __forceinline__ double ppm_g(int i, const double X[4], double dG[4]) {
  const double X0 = X[0], X1 = X[1], X2 = X[2], X3 = X[3];
  dG[0] = dG[1] = dG[2] = dG[3] = 0.0;
  double N, Dn;
  switch (i) {
    case 0: {
      Dn = X0 + X1 + X2;
      const Recip r = recip(Dn);
      const double r2 = r.y * r.y;
      dG[0] = -X1 * r2; dG[1] = (X0 + X2) * r2; dG[2] = -X1 * r2;
      return div_r(X1, r);
    }
    case 1: {
      N = 2.0 * X0 + X1; Dn = X1 + X2;
      const Recip r = recip(Dn);
      const double q = div_r(N, r);
      dG[0] = 2.0 * r.y; dG[1] = r.y - q * r.y; dG[2] = -q * r.y;
      return q;
    }
    case 2: {
      N = X1 + 2.0 * X2; Dn = X0 + X1;
      const Recip r = recip(Dn);
      const double q = div_r(N, r);
      dG[0] = -q * r.y; dG[1] = r.y - q * r.y; dG[2] = 2.0 * r.y;
      return q;
    }
    case 3: {
      Dn = X1 + X2;
      const Recip r = recip(Dn);
      const double r2 = r.y * r.y;
      dG[1] = X2 * r2; dG[2] = -X1 * r2;
      return div_r(X1, r);
    }
    case 4: {
      const double g4 = 1.0 / (X0 + X1 + X2 + X3);
      dG[0] = dG[1] = dG[2] = dG[3] = -g4 * g4;
      return g4;
    }
    case 5: {
      Dn = X1 + X2;
      const Recip r = recip(Dn);
      const double r2 = r.y * r.y;
      dG[1] = 2.0 * X2 * X2 * r2; dG[2] = 2.0 * X1 * X1 * r2;
      return div_r(2.0 * X1 * X2, r);
    }
    case 6: {
      N = X0 + X1; Dn = 2.0 * X1 + X2;
      const Recip r = recip(Dn);
      const double q = div_r(N, r);
      dG[0] = r.y; dG[1] = r.y - 2.0 * q * r.y; dG[2] = -q * r.y;
      return q;
    }
    case 7: {
      N = X3 + X2; Dn = 2.0 * X2 + X1;
      const Recip r = recip(Dn);
      const double q = div_r(N, r);
      dG[1] = -q * r.y; dG[2] = r.y - 2.0 * q * r.y; dG[3] = r.y;
      return q;
    }
    case 8: {
      N = X1 * (X0 + X1); Dn = 2.0 * X1 + X2;
      const Recip r = recip(Dn);
      const double q = div_r(N, r);
      dG[0] = X1 * r.y; dG[1] = (X0 + 2.0 * X1) * r.y - 2.0 * q * r.y; dG[2] = -q * r.y;
      return q;
    }
    default: {
      N = X2 * (X2 + X3); Dn = X1 + 2.0 * X2;
      const Recip r = recip(Dn);
      const double q = div_r(N, r);
      dG[1] = -q * r.y; dG[2] = (2.0 * X2 + X3) * r.y - 2.0 * q * r.y; dG[3] = X2 * r.y;
      return q;
    }
  }
}

// dx(i) of the mirrored ghost layout, i = 0..L+3 is Fortran cell i-1
__device__ __forceinline__ int ghost_src(int i, int L) {
  return i == 0 ? 1 : i == 1 ? 0 : i <= L + 1 ? i - 2 : i == L + 2 ? L - 1 : L - 2;
}

// grid of one column: target thicknesses, partitions, target search
// The target search walks pio(0..L+1) of its column serially; the column's
// pio values are also kept in LDS (pio_s[(L+2) rows][blockDim.x], 16 to 64
// columns per block so they fit in 64 KB), so each search step waits on LDS
// rather than on a global load of the scratch array.
__global__ __launch_bounds__(64) void adj_remap_grid_kernel(AdjRemapArgs a) {
  extern __shared__ double pio_s[];
  const int bs = blockDim.x, tid = threadIdx.x;
  const size_t ncol = (size_t)a.nelem * 16;
  const size_t col = (size_t)blockIdx.x * bs + tid;
  if (col >= ncol) return;
  const int L = a.nlev, ie = (int)(col >> 4), pt = (int)(col & 15);
  const RCol c{&a, col, ncol, (size_t)L + 4};
  auto src = [&](int k) { return a.dp[((size_t)ie * L + k) * 16 + pt]; };
  double ps = 0.0;
  for (int k = 0; k < L; ++k) ps += src(k);
  ps += a.hyai[0] * a.ps0;
  double acc = 0.0, accn = 0.0;
  for (int k = 0; k < L; ++k) {
    const double tg = (a.hyai[k + 1] - a.hyai[k]) * a.ps0 + (a.hybi[k + 1] - a.hybi[k]) * ps;
    c.g(RG_TG, k) = tg;
    c.g(RG_PIO, k) = acc;
    pio_s[k * bs + tid] = acc;
    c.g(RG_PIN, k) = accn;
    acc += src(k);
    accn += tg;
  }
  const double pioL = acc;  // pio(L-1) + dx(L-1)
  c.g(RG_PIO, L) = pioL;
  c.g(RG_PIO, L + 1) = pioL + 1.0;
  c.g(RG_PIN, L) = pioL;
  pio_s[L * bs + tid] = pioL;
  pio_s[(L + 1) * bs + tid] = pioL + 1.0;
  double pin = 0.0;  // pin(k+1), the same running sum as above
  for (int k = 0; k < L; ++k) {
    pin += (a.hyai[k + 1] - a.hyai[k]) * a.ps0 + (a.hybi[k + 1] - a.hybi[k]) * ps;
    const double pk = k + 1 < L ? pin : pioL;
    int kk = k + 1;
    while (kk - 1 <= L + 1 && pio_s[(kk - 1) * bs + tid] <= pk) kk++;
    kk--;
    if (kk == L + 1) kk = L;
    // kid >= 0 for every valid column (pio(0) = 0 <= pin(1)); the clamp keeps
    // the reads in bounds for a column the forward step rejected (NaN)
    const int kid = kk - 1 < 0 ? 0 : kk - 1;
    c.g(RG_KID, k) = (double)kid;
    c.g(RG_Z2, k) = (pk - (pio_s[kid * bs + tid] + pio_s[(kid + 1) * bs + tid]) * 0.5) / src(kid);
  }
}

// one lane = (column, field): forward recompute of the field's PPM
// reconstruction, then the reverse sweep; per-field grid adjoints go to
// scratch for adj_remap_finish_kernel
__global__ __launch_bounds__(256) void adj_remap_field_kernel(AdjRemapArgs a) {
  const size_t ncol = (size_t)a.nelem * 16;
  const size_t lane = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= 3 * ncol) return;
  const int fld = (int)(lane / ncol);
  const size_t col = lane % ncol;
  const int L = a.nlev, ie = (int)(col >> 4), pt = (int)(col & 15);
  const RCol c{&a, col, ncol, (size_t)L + 4};
  auto src = [&](int k) { return a.dp[((size_t)ie * L + k) * 16 + pt]; };
  auto DX = [&](int i) { return src(ghost_src(i, L)); };
  auto fidx = [&](int k) {
    return fld < 2 ? (((size_t)ie * L + k) * 2 + fld) * 16 + pt : ((size_t)ie * L + k) * 16 + pt;
  };
  const double *fin = fld < 2 ? a.v : a.t;
  double *lam = fld < 2 ? a.l_v : a.l_t;
  auto F = [&](int arr, int i) -> double & { return c.f(fld, arr, i); };
  auto A = [&](int j) -> double & { return F(RF_AA, j + 1); };   // a(j), j = -1..L+2
  auto LA = [&](int j) -> double & { return F(RF_LA, j + 1); };
  const int W = L + 4;
  // sums over the targets of a source cell: cells without targets stay zero
  for (int i = 0; i < W; ++i) {
    F(RF_LC0, i) = 0.0; F(RF_LC1, i) = 0.0; F(RF_LC2, i) = 0.0; F(RF_LMASS, i) = 0.0; F(RF_LDX, i) = 0.0;
    F(RF_LPIO, i) = 0.0; F(RF_LPIN, i) = 0.0;
  }
  // ---------------------------------------------------------- forward
  for (int k = 0; k < L; ++k) {
    const double x = src(k);
    A(k + 1) = (fin[fidx(k)] * x) / x;
  }
  A(0) = A(1); A(-1) = A(2); A(L + 1) = A(L); A(L + 2) = A(L - 1);
  double X[4], dG[4];
  for (int j = 0; j < L + 2; ++j) {
    for (int m = 0; m < 4; ++m) X[m] = DX(j + m < L + 3 ? j + m : L + 3);
    const double g0 = ppm_g(0, X, dG), g1 = ppm_g(1, X, dG), g2 = ppm_g(2, X, dG);
    const double dp_ = A(j + 1) - A(j), dm_ = A(j) - A(j - 1);
    double dma = 0.0;
    if (dp_ * dm_ > 0.0) {
      const double da = g0 * (g1 * dp_ + g2 * dm_);
      const double bc = 2.0 * fabs(dm_) < 2.0 * fabs(dp_) ? 2.0 * fabs(dm_) : 2.0 * fabs(dp_);
      dma = (fabs(da) < bc ? fabs(da) : bc) * copysign(1.0, da);
    }
    F(RF_DMA, j) = dma;
  }
  for (int j = 0; j < L + 1; ++j) {
    for (int m = 0; m < 4; ++m) X[m] = DX(j + m);
    double gg[10];
    for (int i = 3; i < 10; ++i) gg[i] = ppm_g(i, X, dG);
    const double d1 = A(j + 1) - A(j);
    F(RF_AI, j) = A(j) + gg[3] * d1 +
                  gg[4] * (gg[5] * (gg[6] - gg[7]) * d1 - gg[8] * F(RF_DMA, j + 1) + gg[9] * F(RF_DMA, j));
  }
  // parabola of cell j (1..L): flags and coefficients
  auto parabola = [&](int j, double &c0, double &c1, double &c2, bool &flat, bool &f2, bool &f3, bool &bnd) {
    const double al0 = F(RF_AI, j - 1), ar0 = F(RF_AI, j), aa = A(j);
    flat = (ar0 - aa) * (aa - al0) <= 0.0;
    double al = flat ? aa : al0, ar = flat ? aa : ar0;
    f2 = (ar - al) * (aa - (al + ar) / 2.0) > (ar - al) * (ar - al) / 6.0;
    if (f2) al = 3.0 * aa - 2.0 * ar;
    f3 = (ar - al) * (aa - (al + ar) / 2.0) < -(ar - al) * (ar - al) / 6.0;
    if (f3) ar = 3.0 * aa - 2.0 * al;
    c0 = 1.5 * aa - (al + ar) / 4.0;
    c1 = ar - al;
    c2 = 3.0 * (-2.0 * aa + (al + ar));
    bnd = a.alg == 2 && (j <= 2 || j >= L - 1);
    if (bnd) { c0 = aa; c1 = 0.0; c2 = 0.0; }
  };
  // The reverse passes accumulate in registers over a sliding window of
  // levels and write each scratch entry once (or add to it once), so no
  // entry is read back while its own update is still in flight.
  // ------------------------------------------ reverse: targets (ascending)
  // kid is non-decreasing in k: the targets of one source cell are
  // consecutive, and the cell's sums are written when the cell changes
  double prefix = 0.0, mprev = 0.0;
  int jj = 0, cur = -1;
  double s_lc0 = 0.0, s_lc1 = 0.0, s_lc2 = 0.0, s_lm = 0.0, s_ldx = 0.0, s_plo = 0.0, s_phi = 0.0;
  auto flush_cell = [&]() {
    if (cur < 0) return;
    F(RF_LC0, cur) = s_lc0;
    F(RF_LC1, cur) = s_lc1;
    F(RF_LC2, cur) = s_lc2;
    F(RF_LMASS, cur) = s_lm;
    F(RF_LDX, cur) = s_ldx;
    F(RF_LPIO, cur) += s_plo;
    F(RF_LPIO, cur + 1) += s_phi;
  };
  double lcur = lam[fidx(0)] / c.g(RG_TG, 0);
  for (int k = 0; k < L; ++k) {
    const int kk = (int)c.g(RG_KID, k);
    if (kk != cur) {
      flush_cell();
      cur = kk;
      s_lc0 = s_lc1 = s_lc2 = s_lm = s_ldx = s_plo = s_phi = 0.0;
    }
    while (jj < kk) { prefix += fin[fidx(jj)] * src(jj); ++jj; }
    const double z2 = c.g(RG_Z2, k), dxk = src(kk), tg = c.g(RG_TG, k);
    double c0, c1, c2; bool fl, f2, f3, bd;
    parabola(kk + 1, c0, c1, c2, fl, f2, f3, bd);
    const double integ = (c0 * (z2 + 0.5) + c1 * (z2 * z2 - 0.25) / 2.0) + c2 * (z2 * z2 * z2 + 0.125) / 3.0;
    const double m = prefix + integ * dxk;
    const double out = k > 0 ? m - mprev : m;
    mprev = m;
    const double lf = lam[fidx(k)];
    F(RF_LTG, k) = -lf * (out / tg) / tg;
    const double lnext = k + 1 < L ? lam[fidx(k + 1)] / c.g(RG_TG, k + 1) : 0.0;
    const double lm = lcur - lnext;
    lcur = lnext;
    const double li = lm * dxk;
    const double lz2 = li * (c0 + c1 * z2 + c2 * z2 * z2);
    s_lm += lm;
    s_ldx += lm * integ - lz2 * z2 / dxk;
    s_lc0 += li * (z2 + 0.5);
    s_lc1 += li * (z2 * z2 - 0.25) / 2.0;
    s_lc2 += li * (z2 * z2 * z2 + 0.125) / 3.0;
    s_plo -= 0.5 * lz2 / dxk;
    s_phi -= 0.5 * lz2 / dxk;
    F(RF_LPIN, k + 1) = lz2 / dxk;
  }
  flush_cell();
  // mass prefix -> lambda of var: LMASS(j) becomes sum_{i > j} LMASS(i)
  {
    double s = F(RF_LMASS, L);
    for (int j = L - 1; j >= 0; --j) {
      const double o = F(RF_LMASS, j);
      F(RF_LMASS, j) = s;
      s += o;
    }
  }
  // cells (descending): coefficients -> a(j), ai(j-1), ai(j); first writer of
  // LAI (all) and LA(1..L)
  LA(-1) = 0.0; LA(0) = 0.0; LA(L + 1) = 0.0; LA(L + 2) = 0.0;
  double lai_carry = 0.0;  // ai(j) as the left edge of cell j+1
  for (int j = L; j >= 1; --j) {
    const int cidx = j - 1;
    double c0, c1, c2; bool flat, f2, f3, bnd;
    parabola(j, c0, c1, c2, flat, f2, f3, bnd);
    const double l0 = F(RF_LC0, cidx), l1 = F(RF_LC1, cidx), l2 = F(RF_LC2, cidx);
    double laa, lal, lar;
    if (bnd) {
      laa = l0; lal = 0.0; lar = 0.0;
    } else {
      laa = 1.5 * l0 - 6.0 * l2;
      lal = -l0 / 4.0 - l1 + 3.0 * l2;
      lar = -l0 / 4.0 + l1 + 3.0 * l2;
      if (f3) { laa += 3.0 * lar; lal += -2.0 * lar; lar = 0.0; }
      if (f2) { laa += 3.0 * lal; lar += -2.0 * lal; lal = 0.0; }
      if (flat) { laa += lal + lar; lal = 0.0; lar = 0.0; }
    }
    F(RF_LAI, j) = lai_carry + lar;
    lai_carry = lal;
    LA(j) = laa;
  }
  F(RF_LAI, 0) = lai_carry;
  // interfaces ai(j), j = L..0; first writer of LDMA and LD
  {
    double la_carry = 0.0, ldma_carry = 0.0, w[4] = {0.0, 0.0, 0.0, 0.0};  // w[m]: LD(j+m)
    for (int j = L; j >= 0; --j) {
      if (j + 4 <= L + 3) F(RF_LD, j + 4) = w[3];
      w[3] = w[2]; w[2] = w[1]; w[1] = w[0]; w[0] = 0.0;
      for (int m = 0; m < 4; ++m) X[m] = DX(j + m);
      double gg[10], dd[10][4];
      for (int i = 3; i < 10; ++i) gg[i] = ppm_g(i, X, dd[i]);
      const double li = F(RF_LAI, j);
      const double d1 = A(j + 1) - A(j);
      const double g56 = gg[5] * (gg[6] - gg[7]);
      const double dmaj = F(RF_DMA, j), dmaj1 = F(RF_DMA, j + 1);
      const double ld1 = li * (gg[3] + gg[4] * g56);
      LA(j + 1) += la_carry + ld1;
      la_carry = li - ld1;
      F(RF_LDMA, j + 1) = ldma_carry - li * gg[4] * gg[8];
      ldma_carry = li * gg[4] * gg[9];
      double lg[10] = {0};
      lg[3] = li * d1;
      lg[4] = li * (g56 * d1 - gg[8] * dmaj1 + gg[9] * dmaj);
      const double lg56 = li * gg[4] * d1;
      lg[5] = lg56 * (gg[6] - gg[7]);
      lg[6] = lg56 * gg[5];
      lg[7] = -lg56 * gg[5];
      lg[8] = -li * gg[4] * dmaj1;
      lg[9] = li * gg[4] * dmaj;
      for (int i = 3; i < 10; ++i)
        for (int m = 0; m < 4; ++m) w[m] += lg[i] * dd[i][m];
    }
    LA(0) += la_carry;
    F(RF_LDMA, 0) = ldma_carry;
    for (int m = 0; m < 4; ++m) F(RF_LD, m) = w[m];
  }
  // dma(j), j = L+1..0
  {
    double c1 = 0.0, c2 = 0.0, w[3] = {0.0, 0.0, 0.0};  // c1: LA(j+1) pending, c2: LA(j); w[m]: LD(j+m)
    for (int j = L + 1; j >= 0; --j) {
      if (j + 3 <= L + 3) F(RF_LD, j + 3) += w[2];
      w[2] = w[1]; w[1] = w[0]; w[0] = 0.0;
      double ldp = 0.0, ldm = 0.0;
      const double dp_ = A(j + 1) - A(j), dm_ = A(j) - A(j - 1);
      if (dp_ * dm_ > 0.0) {
        for (int m = 0; m < 4; ++m) X[m] = DX(j + m < L + 3 ? j + m : L + 3);
        double d0[4], d1[4], d2[4];
        const double g0 = ppm_g(0, X, d0), g1 = ppm_g(1, X, d1), g2 = ppm_g(2, X, d2);
        const double da = g0 * (g1 * dp_ + g2 * dm_);
        const bool bsel = 2.0 * fabs(dm_) < 2.0 * fabs(dp_);
        const double bc = bsel ? 2.0 * fabs(dm_) : 2.0 * fabs(dp_);
        const bool asel = fabs(da) < bc;
        const double sda = da > 0.0 ? 1.0 : da < 0.0 ? -1.0 : 0.0;
        const double lm3 = F(RF_LDMA, j) * sda;
        double lda = 0.0;
        if (asel) {
          lda = lm3 * sda;
        } else if (bsel) {
          ldm = 2.0 * (dm_ > 0.0 ? 1.0 : dm_ < 0.0 ? -1.0 : 0.0) * lm3;
        } else {
          ldp = 2.0 * (dp_ > 0.0 ? 1.0 : dp_ < 0.0 ? -1.0 : 0.0) * lm3;
        }
        const double lg0 = lda * (g1 * dp_ + g2 * dm_), lg1 = lda * g0 * dp_, lg2 = lda * g0 * dm_;
        ldp += lda * g0 * g1;
        ldm += lda * g0 * g2;
        for (int m = 0; m < 3; ++m) w[m] += lg0 * d0[m] + lg1 * d1[m] + lg2 * d2[m];
      }
      LA(j + 1) += c1 + ldp;
      c1 = c2 + (ldm - ldp);
      c2 = -ldm;
    }
    LA(0) += c1;
    LA(-1) += c2;
    for (int m = 0; m < 3; ++m) F(RF_LD, m) += w[m];
  }
  // ghosts of a fold back; a = var/dx, var = f*dx
  for (int k = 0; k < L; ++k) {
    double lao = LA(k + 1);
    if (k == 0) lao += LA(0);
    if (k == 1) lao += LA(-1);
    if (k == L - 1) lao += LA(L + 1);
    if (k == L - 2) lao += LA(L + 2);
    const double x = src(k), fv = fin[fidx(k)], var = fv * x;
    const double lvar = F(RF_LMASS, k) + lao / x;
    F(RF_LDX, k) += -lao * var / (x * x) + lvar * fv;
    lam[fidx(k)] = lvar * x;
  }
}

// per column: the fields' grid adjoints -> adjoint of dp (source thickness)
__global__ __launch_bounds__(256) void adj_remap_finish_kernel(AdjRemapArgs a) {
  const size_t ncol = (size_t)a.nelem * 16;
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int L = a.nlev, ie = (int)(col >> 4), pt = (int)(col & 15);
  const RCol c{&a, col, ncol, (size_t)L + 4};
  auto sum3 = [&](int arr, int i) { return c.f(0, arr, i) + c.f(1, arr, i) + c.f(2, arr, i); };
  // lambda of pio(L) collects pio(L+1) = pio(L) + 1 and pin(L) = pio(L)
  double spio = sum3(RF_LPIO, L) + sum3(RF_LPIO, L + 1) + sum3(RF_LPIN, L);
  double spin = 0.0, lps = a.l_ps[col];
  double *ldp = a.l_dp + (size_t)ie * L * 16 + pt;
  for (int k = L - 1; k >= 0; --k) {
    // dx(k) enters pio(i) for i > k, the ghosts and ps
    double x = sum3(RF_LDX, k) + sum3(RF_LD, k + 2) + spio;
    if (k == 1) x += sum3(RF_LD, 0);
    if (k == 0) x += sum3(RF_LD, 1);
    if (k == L - 1) x += sum3(RF_LD, L + 2);
    if (k == L - 2) x += sum3(RF_LD, L + 3);
    ldp[(size_t)k * 16] = x;
    spio += sum3(RF_LPIO, k);
    // tg(k) enters pin(i) for k < i <= L-1
    const double ltg = sum3(RF_LTG, k) + spin;
    spin += sum3(RF_LPIN, k);
    lps += (a.hybi[k + 1] - a.hybi[k]) * ltg;
  }
  for (int k = 0; k < L; ++k) ldp[(size_t)k * 16] += lps;
}

inline unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_adj_caar(hipStream_t st, const AdjCaarArgs &a) {
  const size_t slab_threads = (size_t)a.nelem * a.nlev * 16;
  const size_t cols = (size_t)a.nelem * 16;
  hipLaunchKernelGGL(adj_caar_div_kernel, dim3(nblk(slab_threads)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_scan_kernel, dim3(nblk(cols)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_b1_kernel, dim3(nblk(slab_threads)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_caar_b2_kernel, dim3(nblk(cols, 64)), dim3(64 * B2C), 0, st, a);
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
void launch_adj_remap(hipStream_t st, const AdjRemapArgs &a) {
  const size_t ncol = (size_t)a.nelem * 16;
  // columns per grid block: as many (up to 64) as keep pio_s within 64 KB
  int gbs = 64;
  while (gbs > 1 && (size_t)(a.nlev + 2) * gbs * sizeof(double) > 65536) gbs >>= 1;
  hipLaunchKernelGGL(adj_remap_grid_kernel, dim3(nblk(ncol, gbs)), dim3(gbs), (size_t)(a.nlev + 2) * gbs * sizeof(double),
                     st, a);
  hipLaunchKernelGGL(adj_remap_field_kernel, dim3(nblk(3 * ncol)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_remap_finish_kernel, dim3(nblk(ncol)), dim3(256), 0, st, a);
}
size_t adj_remap_scratch(int nelem, int nlev) {
  return (size_t)(REMAP_GRID + 3 * REMAP_FIELD) * (nlev + 4) * nelem * 16;
}

void launch_adj_zero(hipStream_t st, size_t n, double *p) {
  hipLaunchKernelGGL(adj_zero_kernel, dim3(nblk(n)), dim3(256), 0, st, n, p);
}

}  // namespace hx
