// Euler tracer step with quasi-monotone limiters (EulerStepFunctorImpl.hpp
// and prim_advec_tracers_remap.cpp in the reference).
//
// Slab kernels (16 lanes = one element at one level) loop over all tracers
// of a slab, so metric, dp and vstar are loaded once per slab.  The exchange
// kernels (DSS, biharmonic second half) are gathers and walk one field at a
// time in level groups (group_slab) so neighbour copies come from L2.  One
// remap interval's tracer step is
//   divdp_qmm  : divdp = div(vn0); stage-1 qmin/qmax
//   nbr_minmax
//   advect(1)  ; dss           (extra field divdp_proj)
//   advect(2)  (stage-2 qmin/qmax update fused: no exchange at this stage)
//   dss        (extra field eta_dot_dpdn)
//   qmm_lap    : stage-3 qmin/qmax and the first laplacian of the biharmonic
//   bih_post   : DSS of the first laplacian, second laplacian, scaling
//   nbr_minmax
//   advect(3)  ; dss + SSP average (extra field omega_p)
// Limiter sums run over the 16 points in index order (slab_sum_serial) to
// round like the reference's serial reductions.
#include <type_traits>

#include "kernels.hpp"

namespace hx {

__device__ __forceinline__ bool slab_ie_k(size_t nslab_total, int nlev, size_t &slab, int &ie, int &k, int &pt) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  slab = (size_t)b * 16 + (threadIdx.x >> 4);
  pt = threadIdx.x & 15;
  if (slab >= nslab_total) return false;
  ie = (int)(slab / nlev);
  k = (int)(slab % nlev);
  return true;
}

// metric needed by divergence / weak laplacian only
__device__ __forceinline__ void load_elem_light(const MetricPtrs &m, int ie, int pt, ElemLane &e) {
  const size_t t = (size_t)ie * 64 + pt, s = (size_t)ie * 16 + pt;
#pragma unroll
  for (int j = 0; j < 4; ++j) e.Dinv[j] = m.Dinv[t + 16 * j];
  e.metdet = m.metdet[s];
  e.sph = m.spheremp[s];
  e.rsmp = m.rspheremp[s];
}

// --------------------------------------------------------------------------
// divdp = div(vn0), divdp_proj = divdp (EulerStepFunctorImpl.hpp:348-377) and
// the stage-1 compute_dp + compute_qmin_qmax (:406-485)
__global__ __launch_bounds__(256) void euler_divdp_qmm_kernel(EulerArgs a) {
  size_t slab;
  int ie, k, pt;
  if (!slab_ie_k((size_t)a.nelem * a.nlev, a.nlev, slab, ie, k, pt)) return;
  const int L = a.nlev, Q = a.qsize;
  ElemLane e;
  load_elem_light(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t s = slab * 16 + pt, v = slab * 32 + pt;
  const double dv = sph_div(e, d, a.vn0[v], a.vn0[v + 16]);
  a.divdp_w[s] = dv;
  a.divdp_proj_w[s] = dv;
  const double rhsmdt = a.rhs_multiplier * a.dt;
  const double dp = a.derived_dp[s] - rhsmdt * dv;
  const Recip rdp = recip(dp);  // Qdp / dp for every tracer
  // the Qdp of the next PD tracers is in flight during a tracer's
  // reductions: slot j of nx holds tracers j, j+PD, j+2PD, ..., refilled as
  // it is consumed (the loop body covers PD tracers, so no value moves
  // between registers: a move would wait for its load).  Every lane stores
  // the slab's bounds (the same value: measured faster here than a store
  // from one lane, whose branch makes the next wait cover the stores).
  if (Q <= 0) return;
  constexpr int PD = 4;
  const size_t q0i = ((size_t)ie * Q * L + k) * 16 + pt, qs = (size_t)L * 16;
  // (past the last tracer the last is reloaded: the loads stay unconditional)
  auto ld = [&](int q) { return a.qdp_n0[q0i + (size_t)(q < Q ? q : Q - 1) * qs]; };
  double nx[PD];
#pragma unroll
  for (int j = 0; j < PD; ++j) nx[j] = ld(j);
  auto tracer = [&](int q, double &slot) {
    const double x = div_r(slot, rdp);
    slot = ld(q + PD);
    const double mn = slab_min(x), mx = slab_max(x);
    double *ql = a.qlim + ((size_t)ie * Q + q) * ql_row(L);
    *(double2 *)(ql + ql_min(k)) = make_double2(mn, mx);
  };
  int q = 0;
  for (; q + PD <= Q; q += PD) {
#pragma unroll
    for (int j = 0; j < PD; ++j) tracer(q + j, nx[j]);
  }
#pragma unroll
  for (int j = 0; j < PD; ++j)
    if (q + j < Q) tracer(q + j, nx[j]);
}

// min/max of qlim over the element and its (up to 8) neighbours (:487-507).
// One thread per (tracer, element, level), tracer slowest, so a neighbour's
// qlim row is still in L2 when the next element row reads it.
__global__ __launch_bounds__(256) void euler_nbr_minmax_kernel(EulerArgs a) {
  // XCD-aware block order: neighbouring elements' rows are read through one L2
  const size_t i = (size_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int L = a.nlev, Q = a.qsize;
  if (i >= (size_t)a.nelem * Q * L) return;
  const int k = (int)(i % L);
  const int ie = (int)((i / L) % a.nelem);
  const int q = (int)(i / ((size_t)L * a.nelem));
  const size_t own = ((size_t)ie * Q + q) * ql_row(L) + ql_min(k);
  const double2 b = *(const double2 *)(a.qlim + own);
  double mn = b.x, mx = b.y;
  // an absent neighbour (-1) reads the element's own row, which leaves the
  // min and max unchanged: no branch, all loads in flight together
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int nb = a.nbr[ie * 8 + c];
    const size_t o = opaque_index(((size_t)(nb < 0 ? ie : nb) * Q + q) * ql_row(L) + ql_min(k));
    const double2 n = *(const double2 *)(a.qlim + o);
    mx = vmax64(mx, n.y);
    mn = vmin64(mn, n.x);
  }
  *(double2 *)(a.qlim_new + own) = make_double2(mn, mx);
}

// gather of point (ie, pt) at level k from pre-exchange fields (layout
// lc_index) with element stride es: the own copy and the three sources'
// copies (an absent source points at the own copy), in the reference's
// order.  The four addresses are formed once per lane; field f of a batch is
// then one offset (f * field size, uniform) added to each.  load issues the
// four loads, gather_sum adds them (FP32 buffers of the mixed form are
// accumulated in FP64 the same way).
template <typename T>
struct LfGather {
  gptr<T> p[4];
  __device__ __forceinline__ LfGather(const T *fg, size_t es, int ie, int pt, int k, int s0, int s1, int s2) {
    const gptr<T> f = (gptr<T>)fg;
    p[0] = f + (size_t)ie * es + lc_index(pt, k);
    // (opaque: the selected address stays one unconditional load)
    auto at = [&](int s) { return opaque_ptr(s >= 0 ? f + (size_t)(s >> 4) * es + lc_index(s & 15, k) : p[0]); };
    p[1] = at(s0);
    p[2] = at(s1);
    p[3] = at(s2);
  }
  __device__ __forceinline__ void load(size_t off, T (&r)[4]) const {
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = p[j][off];
  }
  __device__ __forceinline__ double sum(size_t off, int s0, int s1, int s2) const {
    T r[4];
    load(off, r);
    return gather_sum(r, s0, s1, s2);
  }
};

// The same gather with the addresses formed per field: fewer live registers
// than LfGather's four pointers (euler_dss keeps 8 waves per SIMD with it).
template <typename T>
__device__ __forceinline__ double gather_lf(const T *f, size_t es, size_t off, int ie, int pt, int k, int s0, int s1,
                                            int s2) {
  const size_t own = (size_t)ie * es + off + lc_index(pt, k);
  auto at = [&](int s) { return opaque_index(s >= 0 ? (size_t)(s >> 4) * es + off + lc_index(s & 15, k) : own); };
  const T r[4] = {f[own], f[at(s0)], f[at(s1)], f[at(s2)]};
  return gather_sum(r, s0, s1, s2);
}

// ------------------------------------------------------------ limiters
// limiter_optim_iter_full (:767-823), one slab = one level of one tracer
__device__ __forceinline__ void limit8_slab(double mass, double minp, double maxp, double &x, double c) {
  const double tol_limiter = 5e-14;
  for (int iter = 0; iter < NP2 - 1; ++iter) {
    double delta = 0;
    if (x > maxp) {
      delta = x - maxp;
      x = maxp;
    } else if (x < minp) {
      delta = x - minp;
      x = minp;
    }
    // (skipped when no slab of the wave clips; a slab that clips nothing
    // in a wave that does sums zeros: exactly 0, as if skipped)
    const double addmass = wave_any(delta != 0.0) ? slab_sum_serial(delta * c) : 0.0;
    if (fabs(addmass) <= tol_limiter * fabs(mass)) break;
    if (addmass > 0) {
      const double weightssum = slab_sum_serial(x < maxp ? c : 0.0);
      const double adw = addmass / weightssum;
      x += (x < maxp) ? adw : 0;
    } else {
      const double weightssum = slab_sum_serial(x > minp ? c : 0.0);
      const double adw = addmass / weightssum;
      x += (x > minp) ? adw : 0;
    }
  }
}

// limiter_clip_and_sum (:826-884)
__device__ __forceinline__ void limit9_slab(double minp, double maxp, double &x, double c) {
  double delta = 0;
  if (x > maxp) {
    delta = x - maxp;
    x = maxp;
  } else if (x < minp) {
    delta = x - minp;
    x = minp;
  }
  const double addmass = wave_any(delta != 0.0) ? slab_sum_serial(delta * c) : 0.0;
  if (addmass == 0) return;
  if (addmass > 0) {
    double fac = slab_sum_serial(c * (maxp - x));
    if (fac > 0) {
      fac = addmass / fac;
      x += fac * (maxp - x);
    }
  } else {
    double fac = slab_sum_serial(c * (x - minp));
    if (fac > 0) {
      fac = addmass / fac;
      x += fac * (x - minp);
    }
  }
}

// Setup phase (:585-626) + tracer phase (:571-687) for every tracer of an
// (element, level) slab.  At stage 2 the qmin/qmax update of that stage
// (min/max with the previous bounds, no neighbour exchange) runs first.
//
// GATHER (stage 2 only): the stage-1 exchange is done here instead of by a
// separate kernel.  A block is one element x AGL levels (16*AGL lanes); lanes
// (pt, kk) gather the stage-1 pre-exchange values (levels fastest) of point
// pt at level k0+kk with the reference's accumulation order, an LDS tile
// hands them to the slab lanes (kk, pt), and the assembled stage-1 Qdp is
// used directly (never stored).  The assembled extra field (divdp_proj) is
// stored, later stages read it.
constexpr int AGL = 8;

// (at most 96 VGPRs: five waves per SIMD)
template <bool MIXED, bool GATHER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void euler_advect_kernel(EulerArgs a,
                                                                                                     int stage) {
  __shared__ double tile[GATHER ? 2 : 1][AGL][17];
  size_t slab;
  int ie, k, pt;
  const int L = a.nlev, Q = a.qsize;
  // gather role (GATHER only)
  int gpt = 0, gk = 0, gkk = 0, kk = 0, g0 = -1, g1 = -1, g2 = -1;
  bool gon = false, son = true;
  double gr = 0.0;
  if (GATHER) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    ie = lb % a.nelem;
    const int k0 = (lb / a.nelem) * AGL;
    if (k0 >= L) return;  // whole block
    kk = threadIdx.x >> 4;
    pt = threadIdx.x & 15;
    k = k0 + kk;
    son = k < L;
    gpt = threadIdx.x / AGL;
    gkk = threadIdx.x % AGL;
    gk = k0 + gkk;
    gon = gk < L;
    const int *src = a.src + ((size_t)ie * 16 + gpt) * 3;
    g0 = src[0];
    g1 = src[1];
    g2 = src[2];
    gr = a.m.rspheremp[(size_t)ie * 16 + gpt];
    slab = (size_t)ie * L + (son ? k : 0);
  } else {
    if (!slab_ie_k((size_t)a.nelem * a.nlev, a.nlev, slab, ie, k, pt)) return;
  }
  ElemLane e;
  load_elem_light(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t s = slab * 16 + pt, v = slab * 32 + pt;
  const double rhsmdt = a.rhs_multiplier * a.dt;
  double dpproj;
  const int gkc = gon ? gk : L - 1;  // gathers of a padding lane stay in bounds
  if (GATHER) {
    // assembled divdp_proj of stage 1
    const double y = LfGather<double>(a.fdss_pre_in, lc_field(L), ie, gpt, gkc, g0, g1, g2).sum(0, g0, g1, g2) * gr;
    if (gon) tile[0][gkk][gpt] = y;
    lds_barrier();
    dpproj = tile[0][kk][pt];
    if (son) a.divdp_proj_w[s] = dpproj;
    lds_barrier();
  } else {
    dpproj = a.divdp_proj[s];
  }
  const double dp = a.derived_dp[s] - rhsmdt * dpproj;
  const double vs0 = a.vn0[v] / dp, vs1 = a.vn0[v + 16] / dp;
  double dpdissk = dp - a.dt * a.divdp[s];
  // (the gathering kernel runs stage 2 only, where rhs_viss == 0: no
  // biharmonic terms, and no conditional loads that would wait for the gathers)
  const bool add_ps_diss = !GATHER && a.nu_p > 0 && a.rhs_viss != 0.0;
  if (add_ps_diss) {
    const double diss_fac = -a.rhs_viss * a.dt * a.nu_q;
    dpdissk += diss_fac * a.dpdiss_bih[s] / e.sph;
  }
  if (son) a.fdss_pre[(size_t)ie * lc_field(L) + lc_index(pt, k)] = a.fdss[s] * e.sph;
  const bool add_hv = !GATHER && a.rhs_viss != 0.0;
  const bool limit = a.limiter_option == 8 || a.limiter_option == 9;
  DvvLaneF df;
  float rmd = 0.f;
  if (MIXED) {
    load_dvvf(a.m.dvv, pt, df);
    rmd = (float)(1.0 / e.metdet * RREARTH);
  }
  const double c = e.sph * dpdissk;
  // the divisors shared by the tracers: Qdp / dp (stage 2), ptens / dpdissk
  const Recip rdp = recip(dp), rdpd = recip(dpdissk);
  const double s1 = limit ? slab_sum_serial(c) : 0.0;  // same for every tracer
  const size_t fs = lc_field(L), es = (size_t)Q * fs;
  // the tracer pre-exchange values are stored in FP32 in the mixed form (the
  // exchange sums them in FP64)
  using PreT = typename std::conditional<MIXED, float, double>::type;
  const PreT *qpin = (const PreT *)a.qdp_pre_in;
  PreT *qpout = (PreT *)a.qdp_pre;
  PreT graw[4] = {};  // the loads of the next tracer's gather
  const LfGather<PreT> gq(qpin, es, ie, gpt, gkc, g0, g1, g2);
  if (GATHER && Q > 0) gq.load(0, graw);
  const int kc = son ? k : L - 1;  // reads of a padding slab stay in bounds
  // Slab path (stages 1 and 3): tracer q+1's Qdp, biharmonic term and
  // bounds are loaded during tracer q's arithmetic.  The loads are
  // unconditional (past the last tracer the last is reloaded; without a
  // biharmonic term of a type, Qdp, or one FP32 value, is loaded in its
  // place): a conditional load would make the waits that follow cover the
  // prefetch.
  constexpr bool PF = !GATHER;
  const bool bihf = add_hv && a.bih32;  // FP32 biharmonic term (mixed form)
  const double *qbp = add_hv && !a.bih32 ? a.qbih : a.qdp_n0;
  const float *qbf = (const float *)(bihf ? a.qbih : a.qdp_n0);
  double pf[5] = {};  // Qdp, biharmonic term (FP64, FP32), qmin, qmax of the next tracer
  // the lane's addresses for tracer 0; tracer q adds q times a (uniform)
  // tracer stride
  const size_t qs = (size_t)L * 16, q0i = ((size_t)ie * Q * L + kc) * 16 + pt;
  const double *qdp_l = a.qdp_n0 + q0i, *qbp_l = qbp + q0i;
  const float *qbf_l = bihf ? qbf + q0i : qbf;
  double *ql_l = a.qlim + (size_t)ie * Q * ql_row(L) + ql_min(kc);
  PreT *qpo_l = qpout + (size_t)ie * es + lc_index(pt, kc);
  auto fetch = [&](int q) {
    const size_t qq = q < Q ? q : Q - 1;
    const double2 b = *(const double2 *)(ql_l + qq * ql_row(L));
    pf[2] = b.x;
    pf[3] = b.y;
    pf[0] = qdp_l[qq * qs];
    pf[1] = qbp_l[qq * qs];
    if (MIXED) pf[4] = qbf_l[bihf ? qq * qs : 0];
  };
  if (PF && Q > 0) fetch(0);
  // tracer q's work; at stage 2 gcur holds tracer q's gathered loads and
  // takes tracer q+GD's.  The mixed form (FP32 gathers, 96 VGPRs) keeps two
  // tracers of gathers in flight (1581 against 1664 us); FP64 one (two need
  // 112 VGPRs, 4 waves per SIMD: 1820 against 1708 us).
  constexpr int GD = MIXED ? 2 : 1;
  auto tracer = [&](int q, PreT (&gcur)[4]) {
    const size_t qi = (size_t)q * qs;  // tracer q's offset from qdp_l, qbp_l, qbf_l
    double *ql = ql_l + (size_t)q * ql_row(L);
    double minp, maxp, qd, q0pf = 0.0;
    if (PF) {
      qd = pf[0];
      q0pf = add_hv ? (bihf ? pf[4] : pf[1]) : 0.0;
      minp = pf[2];
      maxp = pf[3];
      fetch(q + 1);
    } else {
      // this tracer's bounds are loaded before the next tracer's gather is
      // issued: loads complete in issue order, so waiting for them does not
      // wait for the gather
      minp = ql[0];
      maxp = ql[QL_MAX];
    }
    if (GATHER) {
      // the stage-1 exchange of tracer q; tracer q+GD's loads are issued
      // before this tracer's arithmetic and stay in flight across the
      // (LDS-only) barrier
      const double y = gather_sum(gcur, g0, g1, g2) * gr;
      if (gon) tile[q & 1][gkk][gpt] = y;
      // (unconditional, the last tracer reloads itself: a branch here would
      // make every later wait cover the gathers)
      gq.load((size_t)(q + GD < Q ? q + GD : Q - 1) * fs, gcur);
      lds_barrier();
      qd = tile[q & 1][kk][pt];
    } else if (!PF) {
      qd = qdp_l[qi];
    }
    if (stage == 2) {  // compute_qmin_qmax with rhs_multiplier == 1
      const double xq = div_r(qd, rdp);
      minp = dmin(minp, slab_min(xq));
      maxp = dmax(maxp, slab_max(xq));
    }
    double ptens;
    if (MIXED) {
      // flux divergence in FP32 with compensated contractions; the update
      // qdp - dt*div + qbih stays FP64
      const double q0 = PF ? q0pf : (add_hv ? (bihf ? (double)qbf_l[qi] : qbp_l[qi]) : 0);
      const double fa = vs0 * qd, fb = vs1 * qd;
      const float gv0 = (float)((e.Dinv[0] * fa + e.Dinv[2] * fb) * e.metdet);
      const float gv1 = (float)((e.Dinv[1] * fa + e.Dinv[3] * fb) * e.metdet);
      const float dv = (row_contractf(df.J, gv0) + col_contractf(df.I, gv1)) * rmd;
      ptens = (qd + -a.dt * (double)dv + q0);
    } else {
      ptens = sph_div_update(e, d, -a.dt, vs0, vs1, qd, q0pf);
    }
    bool wmin = stage == 2, wmax = stage == 2;
    if (limit) {
      double x = div_r(ptens, rdpd);
      const double s0 = slab_sum_serial(x * c);
      if (!(s1 <= 0)) {
        if (minp < 0) {
          minp = 0;
          wmin = true;
        }
        if (s0 < minp * s1) {
          minp = s0 / s1;
          wmin = true;
        }
        if (s0 > maxp * s1) {
          maxp = s0 / s1;
          wmax = true;
        }
        if (a.limiter_option == 8) limit8_slab(s0, minp, maxp, x, c);
        else limit9_slab(minp, maxp, x, c);
        ptens = x * dpdissk;
      }
    }
    if (pt == 0 && son) {
      if (wmin) ql[0] = minp;
      if (wmax) ql[QL_MAX] = maxp;
    }
    if (son) qpo_l[(size_t)q * fs] = (PreT)(e.sph * ptens);
  };
  if (GATHER && GD == 2) {
    PreT gnext[4] = {};
    if (Q > 1) gq.load(fs, gnext);
    int q = 0;
    for (; q + 1 < Q; q += 2) {
      tracer(q, graw);
      tracer(q + 1, gnext);
    }
    if (q < Q) tracer(q, graw);
  } else {
    for (int q = 0; q < Q; ++q) tracer(q, graw);
  }
}

// ---------------------------------------------------------------- exchanges
// The pre-exchange buffers (qdp_pre, qbih_pre, fdss_pre) are stored levels
// fastest in chunks of 8 levels (lc_index), so a neighbour copy of the 8
// levels of one GLL point in a chunk is one 64-byte run, half of a line whose
// other half the same block reads.  An exchange block is one element x one
// 8-level chunk (XL; 128 lanes, no level padding when the level count is a
// multiple of 8): lane (pt, kk) gathers point pt at level k0+kk, then the
// assembled values go through an LDS tile to lanes (kk, pt) and are stored in
// the slab layout [ie][q][lev][pt].  Blocks walk the elements of one level
// group in order, so neighbour runs come from L2.
constexpr int XL = 8;
constexpr int QB = 8;  // tracers per exchange block

inline int xchg_blocks_x(int nelem, int nlev) {
  const int nb = nelem * ((nlev + XL - 1) / XL);
  return (nb + 7) & ~7;
}

struct XLane {
  int ie, k0, pt, kk, k;  // gather role: point pt, level k = k0 + kk
  int wk, wpt;            // store role: level k0 + wk, point wpt
};

__device__ __forceinline__ bool xchg_lane(int nelem, int nlev, XLane &x) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  x.ie = lb % nelem;
  const int g = lb / nelem;
  x.k0 = g * XL;
  if (x.k0 >= nlev) return false;  // whole block
  x.pt = threadIdx.x / XL;
  x.kk = threadIdx.x % XL;
  x.k = x.k0 + x.kk;
  x.wk = threadIdx.x >> 4;
  x.wpt = threadIdx.x & 15;
  return true;
}


// DSS of tracers [QB*blockIdx.y, QB*blockIdx.y + QB) and, in the first batch,
// of the extra field.  avg: stage 3's SSP average (Qdp(n0) + 2 Qdp(np1)) / 3
// (:379-403) of the assembled tracer value.  F32: the tracer pre-exchange
// values are FP32 (mixed form), summed in FP64.
// QMM (stage 2 only): the same block then does euler_qmm_lap_kernel's work on
// the assembled values in LDS -- the stage-3 bounds (compute_qmin_qmax with
// rhs_multiplier 2, a.rhs_multiplier set to 2) and the first laplacian of the
// tracer biharmonic into qbih_pre -- so Qdp(np1) is not read back.
template <bool F32, bool QMM>
__global__ __launch_bounds__(256) void euler_dss_kernel(EulerArgs a, int avg) {
  // every assembled field of the block goes into LDS before one barrier
  __shared__ double tile[QB + 1][XL][17];
  XLane x;
  if (!xchg_lane(a.nelem, a.nlev, x)) return;
  const int L = a.nlev, Q = a.qsize, q0 = blockIdx.y * QB;
  const int *src = a.src + ((size_t)x.ie * 16 + x.pt) * 3;
  const int s0 = src[0], s1 = src[1], s2 = src[2];
  const double r = a.m.rspheremp[(size_t)x.ie * 16 + x.pt];
  const bool gk = x.k < L, wkv = x.k0 + x.wk < L;
  const size_t fs = lc_field(L), es = (size_t)Q * fs;
  const int nq = (Q - q0) < QB ? (Q - q0) : QB;
  const int kc = gk ? x.k : L - 1;  // gathers of a padding lane stay in bounds
  // every load of the block's gathers is issued before the first sum
  double v[QB], fr[4];
  using PreT = typename std::conditional<F32, float, double>::type;
  if (nq > 0) {
#pragma unroll
    for (int i = 0; i < QB; ++i)
      v[i] = gather_lf((const PreT *)a.qdp_pre, es, (size_t)(q0 + (i < nq ? i : 0)) * fs, x.ie, x.pt, kc, s0, s1, s2);
  }
  if (q0 == 0) LfGather<double>(a.fdss_pre, fs, x.ie, x.pt, kc, s0, s1, s2).load(0, fr);
  if (gk) {
#pragma unroll
    for (int i = 0; i < QB; ++i)
      if (i < nq) tile[i][x.kk][x.pt] = v[i] * r;
    if (q0 == 0) tile[QB][x.kk][x.pt] = gather_sum(fr, s0, s1, s2) * r;
  }
  // stage 3: the Qdp(n0) values of the average are loaded before the barrier
  const size_t o0 = (((size_t)x.ie * Q + q0) * L + x.k0 + x.wk) * 16 + x.wpt, os = (size_t)L * 16;
  double av[QB];
  if (avg && wkv && nq > 0) {
#pragma unroll
    for (int i = 0; i < QB; ++i) av[i] = a.qdp_avg_n0[o0 + (size_t)(i < nq ? i : 0) * os];
  }
  lds_barrier();
  if (!QMM) {
    if (wkv) {
#pragma unroll
      for (int i = 0; i < QB; ++i) {
        if (i >= nq) break;
        double y = tile[i][x.wk][x.wpt];
        if (avg) y = div_c<3>(av[i] + (3.0 - 1) * y);
        st_stream(&a.qdp_np1[o0 + (size_t)i * os], y);
      }
      if (q0 == 0) st_stream(&a.fdss[((size_t)x.ie * L + x.k0 + x.wk) * 16 + x.wpt], tile[QB][x.wk][x.wpt]);
    }
    return;
  }
  // stage-3 bounds and first laplacian (euler_qmm_lap_kernel's operations);
  // every lane of a slab takes part in the cross-lane reductions
  const int wk = x.k0 + x.wk, wkc = wkv ? wk : L - 1;
  const size_t sl = ((size_t)x.ie * L + wkc) * 16 + x.wpt;
  const double rhsmdt = a.rhs_multiplier * a.dt;
  const double dp = a.derived_dp[sl] - rhsmdt * a.divdp_proj[sl];
  const double wdp = a.dpdiss_ave[sl];
  const Recip rdp = recip(dp), rdp0 = recip(a.dp0[wkc]);
  const bool scale = a.nu_p > 0;
  ElemLane e;
  DvvLane d;
  ElemLaneF ef;
  DvvLaneF df;
  if (F32) {
    load_elemf(a.m, x.ie, x.wpt, ef);
    load_dvvf(a.m.dvv, x.wpt, df);
  } else {
    load_elem_light(a.m, x.ie, x.wpt, e);
    load_dvv(a.m.dvv, x.wpt, d);
  }
  double *ql = a.qlim + ((size_t)x.ie * Q + q0) * ql_row(L) + ql_min(wkc);
  const size_t ob = ((size_t)x.ie * Q + q0) * fs + lc_index(x.wpt, wkc);
#pragma unroll
  for (int i = 0; i < QB; ++i) {
    if (i >= nq) break;
    const double y = wkv ? tile[i][x.wk][x.wpt] : 0.0;
    if (wkv) st_stream(&a.qdp_np1[o0 + (size_t)i * os], y);
    const double xq = div_r(y, rdp);
    const double mn = slab_min(xq), mx = slab_max(xq);
    if (wkv && x.wpt == 0) *(double2 *)(ql + (size_t)i * ql_row(L)) = make_double2(mn, mx);
    const double yl = scale ? div_r(xq * wdp, rdp0) : xq;
    if (F32) {
      const float lap = sph_laplacef(ef, df, (float)yl);
      if (wkv) ((float *)a.qbih_pre)[ob + (size_t)i * fs] = lap;
    } else {
      const double lap = sph_laplace(e, d, yl);
      if (wkv) a.qbih_pre[ob + (size_t)i * fs] = lap;
    }
  }
  if (q0 == 0 && wkv) st_stream(&a.fdss[((size_t)x.ie * L + x.k0 + x.wk) * 16 + x.wpt], tile[QB][x.wk][x.wpt]);
}

// Stage 3 compute_dp + compute_qmin_qmax (rhs_multiplier 2, :406-485) and the
// first half of the tracer biharmonic (:196-267): laplace(Q * dpdiss_ave / dp0)
// for every tracer of an (element, level) slab.
template <bool MIXED>
__global__ __launch_bounds__(256) void euler_qmm_lap_kernel(EulerArgs a) {
  size_t slab;
  int ie, k, pt;
  if (!slab_ie_k((size_t)a.nelem * a.nlev, a.nlev, slab, ie, k, pt)) return;
  const int L = a.nlev, Q = a.qsize;
  ElemLane e;
  load_elem_light(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t s = slab * 16 + pt;
  const double rhsmdt = a.rhs_multiplier * a.dt;
  const double dp = a.derived_dp[s] - rhsmdt * a.divdp_proj[s];
  const double wdp = a.dpdiss_ave[s];
  const double dp0 = a.dp0[k];
  const Recip rdp = recip(dp), rdp0 = recip(dp0);  // shared by the tracers
  ElemLaneF ef;
  DvvLaneF df;
  if (MIXED) {
    load_elemf(a.m, ie, pt, ef);
    load_dvvf(a.m.dvv, pt, df);
  }
  if (Q <= 0) return;
  // tracer q+1's Qdp is loaded before tracer q's arithmetic
  const size_t q0i = ((size_t)ie * Q * L + k) * 16 + pt, qs = (size_t)L * 16;
  const bool scale = a.nu_p > 0;
  double nx = a.qdp_n0[q0i];
  auto tracer = [&](int q) {
    const double x = div_r(nx, rdp);
    nx = a.qdp_n0[q0i + (size_t)(q + 1 < Q ? q + 1 : q) * qs];
    const double mn = slab_min(x), mx = slab_max(x);
    if (pt == 0) {
      double *ql = a.qlim + ((size_t)ie * Q + q) * ql_row(L);
      *(double2 *)(ql + ql_min(k)) = make_double2(mn, mx);
    }
    const double y = scale ? div_r(x * wdp, rdp0) : x;
    const size_t o = ((size_t)ie * Q + q) * lc_field(L) + lc_index(pt, k);
    if (MIXED) ((float *)a.qbih_pre)[o] = sph_laplacef(ef, df, (float)y);
    else a.qbih_pre[o] = sph_laplace(e, d, y);
  };
  // the first tracer is peeled so that on both paths into the loop the same
  // stores follow the prefetch (the wait then covers the prefetch only)
  tracer(0);
  for (int q = 1; q < Q; ++q) tracer(q);
}

// biharmonic, second half (:216-310): DSS of the first laplacian (with
// rspheremp), second laplacian, times -rhs_viss*dt*nu_q*dp0/spheremp, for
// tracers [QB*blockIdx.y, QB*blockIdx.y + QB).  After the LDS transpose the
// 16-lane groups are slabs (one level, 16 points) for the laplacian.
template <bool MIXED, bool TENSOR>
__global__ __launch_bounds__(256) void euler_bih_post_kernel(EulerArgs a) {
  __shared__ double tile[QB][XL][17];  // every tracer of the block, one barrier
  XLane x;
  if (!xchg_lane(a.nelem, a.nlev, x)) return;
  const int L = a.nlev, Q = a.qsize, q0 = blockIdx.y * QB;
  const int *src = a.src + ((size_t)x.ie * 16 + x.pt) * 3;
  const int s0 = src[0], s1 = src[1], s2 = src[2];
  const bool gk = x.k < L;
  const int wk = x.k0 + x.wk;
  const bool wkv = wk < L;
  const size_t fs = lc_field(L), es = (size_t)Q * fs;
  const int nq = (Q - q0) < QB ? (Q - q0) : QB;
  const int kc = gk ? x.k : L - 1;  // gathers of a padding lane stay in bounds
  // every load of the block's gathers is issued before the first sum
  double v[QB];
  {
    using PreT = typename std::conditional<MIXED, float, double>::type;
    const LfGather<PreT> g((const PreT *)a.qbih_pre, es, x.ie, x.pt, kc, s0, s1, s2);
#pragma unroll
    for (int i = 0; i < QB; ++i) v[i] = g.sum((size_t)(q0 + (i < nq ? i : 0)) * fs, s0, s1, s2);
  }
  if (gk) {
#pragma unroll
    for (int i = 0; i < QB; ++i)
      if (i < nq) tile[i][x.kk][x.pt] = v[i];
  }
  lds_barrier();
  // the metric is loaded after the barrier, which the gathers' loads do not
  // cross: loaded before it, it would be live during them (69 VGPRs instead
  // of 120 for the FP64 form)
  ElemLane e;
  load_elem_light(a.m, x.ie, x.wpt, e);
  DvvLane d;
  load_dvv(a.m.dvv, x.wpt, d);
  ElemLaneF ef;
  DvvLaneF df;
  if (MIXED) {
    load_elemf(a.m, x.ie, x.wpt, ef);
    load_dvvf(a.m.dvv, x.wpt, df);
  }
  double tv[4], vs[6];
  if (TENSOR) load_tensor(a.m, x.ie, x.wpt, tv, vs);
  const double f = -a.rhs_viss * a.dt * a.nu_q;
  const double dp0 = wkv ? a.dp0[wk] : 0.0;
  const Recip rsph = recip(e.sph);  // shared by the tracers
#pragma unroll
  for (int i = 0; i < QB; ++i) {
    if (i >= nq) break;
    // every lane of a slab takes part in the laplacian's cross-lane reads
    const double y = wkv ? tile[i][x.wk][x.wpt] * e.rsmp : 0.0;
    const size_t o = (((size_t)x.ie * Q + q0 + i) * L + wk) * 16 + x.wpt;
    if (MIXED) {
      const double lap = sph_laplacef(ef, df, (float)y);
      if (wkv) st_stream(&((float *)a.qbih)[o], (float)div_r(f * dp0 * lap, rsph));
    } else if (TENSOR) {
      // BIHPostTensorHV (EulerStepFunctorImpl.hpp:281-291)
      const double lap = sph_laplace_tensor(e, d, tv, y);
      if (wkv) st_stream(&a.qbih[o], div_r(f * dp0 * lap, rsph));
    } else {
      const double lap = sph_laplace(e, d, y);
      if (wkv) st_stream(&a.qbih[o], div_r(f * dp0 * lap, rsph));
    }
  }
}

static int slab_blocks(const EulerArgs &a) { return (int)(((size_t)a.nelem * a.nlev + 15) / 16); }

void launch_euler_divdp_qmm(hipStream_t st, const EulerArgs &a) {
  hipLaunchKernelGGL(euler_divdp_qmm_kernel, dim3(slab_blocks(a)), dim3(256), 0, st, a);
}

void launch_euler_nbr_minmax(hipStream_t st, const EulerArgs &a) {
  const size_t n = (size_t)a.nelem * a.qsize * a.nlev;
  hipLaunchKernelGGL(euler_nbr_minmax_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_euler_advect(hipStream_t st, const EulerArgs &a, int stage) {
  if (a.mixed) hipLaunchKernelGGL((euler_advect_kernel<true, false>), dim3(slab_blocks(a)), dim3(256), 0, st, a, stage);
  else hipLaunchKernelGGL((euler_advect_kernel<false, false>), dim3(slab_blocks(a)), dim3(256), 0, st, a, stage);
}

void launch_euler_advect_gather(hipStream_t st, const EulerArgs &a) {
  const int nb = ((a.nelem * ((a.nlev + AGL - 1) / AGL)) + 7) & ~7;
  if (a.mixed) hipLaunchKernelGGL((euler_advect_kernel<true, true>), dim3(nb), dim3(16 * AGL), 0, st, a, 2);
  else hipLaunchKernelGGL((euler_advect_kernel<false, true>), dim3(nb), dim3(16 * AGL), 0, st, a, 2);
}

void launch_euler_dss(hipStream_t st, const EulerArgs &a, int avg) {
  const int nq = a.qsize > 0 ? (a.qsize + QB - 1) / QB : 1;
  const dim3 g(xchg_blocks_x(a.nelem, a.nlev), nq);
  if (a.mixed) hipLaunchKernelGGL((euler_dss_kernel<true, false>), g, dim3(16 * XL), 0, st, a, avg);
  else hipLaunchKernelGGL((euler_dss_kernel<false, false>), g, dim3(16 * XL), 0, st, a, avg);
}

bool launch_euler_dss_qmm(hipStream_t st, const EulerArgs &a) {
  // Used in the mixed form (FP32 pre-exchange values and FP32 biharmonic
  // buffers), where it measured 1.72 ms against 1.90 for the two launches.
  // In FP64 the fused kernel (88 VGPRs, 5 waves per SIMD against 7) measured
  // 2.13 ms against 2.08, so FP64 keeps two launches, as does the mixed form
  // with tensor hyperviscosity (FP64 biharmonic buffers).
  if (a.qsize <= 0 || !a.mixed || !a.bih32) return false;
  const dim3 g(xchg_blocks_x(a.nelem, a.nlev), (a.qsize + QB - 1) / QB);
  hipLaunchKernelGGL((euler_dss_kernel<true, true>), g, dim3(16 * XL), 0, st, a, 0);
  return true;
}

void launch_euler_qmm_lap(hipStream_t st, const EulerArgs &a) {
  if (a.bih32) hipLaunchKernelGGL(euler_qmm_lap_kernel<true>, dim3(slab_blocks(a)), dim3(256), 0, st, a);
  else hipLaunchKernelGGL(euler_qmm_lap_kernel<false>, dim3(slab_blocks(a)), dim3(256), 0, st, a);
}

void launch_euler_bih_post(hipStream_t st, const EulerArgs &a) {
  if (a.qsize <= 0) return;
  const dim3 g(xchg_blocks_x(a.nelem, a.nlev), (a.qsize + QB - 1) / QB);
  if (a.bih32) hipLaunchKernelGGL((euler_bih_post_kernel<true, false>), g, dim3(16 * XL), 0, st, a);
  else if (a.consthv) hipLaunchKernelGGL((euler_bih_post_kernel<false, false>), g, dim3(16 * XL), 0, st, a);
  else hipLaunchKernelGGL((euler_bih_post_kernel<false, true>), g, dim3(16 * XL), 0, st, a);
}

}  // namespace hx
