// Euler tracer step with quasi-monotone limiters (EulerStepFunctorImpl.hpp
// in the reference).  Slab kernels (16 lanes = one element at one level).
// The advection kernel keeps one slab per 16 lanes and loops over all
// tracers, so the per-level inputs (vstar, dpdissk, metric) are loaded once
// and reused qsize times.  Limiter sums are serial over the 16 points
// (slab_sum_serial) to round like the reference.
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

// divdp = div(vn0), divdp_proj = divdp (EulerStepFunctorImpl.hpp:348-377)
__global__ __launch_bounds__(256) void euler_divdp_kernel(EulerArgs a) {
  size_t slab;
  int ie, k, pt;
  if (!slab_ie_k((size_t)a.nelem * a.nlev, a.nlev, slab, ie, k, pt)) return;
  ElemLane e;
  load_elem(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t s = slab * 16 + pt, v = slab * 32 + pt;
  const double dv = sph_div(e, d, a.vn0[v], a.vn0[v + 16]);
  a.divdp_w[s] = dv;
  a.divdp_proj_w[s] = dv;
}

void launch_euler_divdp(hipStream_t st, const EulerArgs &a) {
  const size_t nslab = (size_t)a.nelem * a.nlev;
  hipLaunchKernelGGL(euler_divdp_kernel, dim3((nslab + 15) / 16), dim3(256), 0, st, a);
}

// compute_dp + compute_qmin_qmax (:406-485).  Slab = (ie, q, lev).
__global__ __launch_bounds__(256) void euler_qminmax_kernel(EulerArgs a, int write_q) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const size_t slab = (size_t)b * 16 + (threadIdx.x >> 4);
  const int pt = threadIdx.x & 15;
  const int L = a.nlev, Q = a.qsize;
  if (slab >= (size_t)a.nelem * Q * L) return;
  const int k = (int)(slab % L);
  const int q = (int)((slab / L) % Q);
  const int ie = (int)(slab / ((size_t)L * Q));
  const size_t s = ((size_t)ie * L + k) * 16 + pt;
  const double rhsmdt = a.rhs_multiplier * a.dt;
  const double dp = a.derived_dp[s] - rhsmdt * a.divdp_proj[s];
  const size_t qi = slab * 16 + pt;  // [ie][q][lev][pt]
  const double v = a.qdp_n0[qi] / dp;
  if (write_q) a.qbih[qi] = v;
  const double mn = slab_min(v), mx = slab_max(v);
  if (pt == 0) {
    double *ql = a.qlim + (((size_t)ie * Q + q) * 2) * L;
    if (a.rhs_multiplier != 1.0) {
      ql[k] = mn;
      ql[L + k] = mx;
    } else {
      ql[k] = dmin(ql[k], mn);
      ql[L + k] = dmax(ql[L + k], mx);
    }
  }
}

// min/max of qlim over the element and its (up to 8) neighbours (:487-507)
__global__ __launch_bounds__(256) void euler_nbr_minmax_kernel(EulerArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int L = a.nlev, Q = a.qsize;
  if (i >= (size_t)a.nelem * Q * L) return;
  const int k = (int)(i % L);
  const int q = (int)((i / L) % Q);
  const int ie = (int)(i / ((size_t)L * Q));
  const size_t own = (((size_t)ie * Q + q) * 2) * L + k;
  double mn = a.qlim[own], mx = a.qlim[own + L];
  for (int c = 0; c < 8; ++c) {
    const int nb = a.nbr[ie * 8 + c];
    if (nb < 0) continue;
    const size_t o = (((size_t)nb * Q + q) * 2) * L + k;
    mx = dmax(mx, a.qlim[o + L]);
    mn = dmin(mn, a.qlim[o]);
  }
  a.qlim_new[own] = mn;
  a.qlim_new[own + L] = mx;
}

// biharmonic pre (:196-267): qbih <- laplace(qbih * dpdiss_ave / dp0)
__global__ __launch_bounds__(256) void euler_bih_pre_kernel(EulerArgs a) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const size_t slab = (size_t)b * 16 + (threadIdx.x >> 4);
  const int pt = threadIdx.x & 15;
  const int L = a.nlev, Q = a.qsize;
  if (slab >= (size_t)a.nelem * Q * L) return;
  const int k = (int)(slab % L);
  const int ie = (int)(slab / ((size_t)L * Q));
  ElemLane e;
  load_elem(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t qi = slab * 16 + pt;
  double x = a.qbih[qi];
  if (a.nu_p > 0) x = x * a.dpdiss_ave[((size_t)ie * L + k) * 16 + pt] / a.dp0[k];
  a.qbih_pre[qi] = sph_laplace(e, d, x);
}

// biharmonic post (:216-310): laplace of the DSS'd field, times -rhs_viss*dt*nu_q*dp0/spheremp
__global__ __launch_bounds__(256) void euler_bih_post_kernel(EulerArgs a) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const size_t slab = (size_t)b * 16 + (threadIdx.x >> 4);
  const int pt = threadIdx.x & 15;
  const int L = a.nlev, Q = a.qsize;
  if (slab >= (size_t)a.nelem * Q * L) return;
  const int k = (int)(slab % L);
  const int ie = (int)(slab / ((size_t)L * Q));
  ElemLane e;
  load_elem(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t qi = slab * 16 + pt;
  const double lap = sph_laplace(e, d, a.qbih[qi]);
  const double f = -a.rhs_viss * a.dt * a.nu_q;
  a.qbih[qi] = (f * a.dp0[k] * lap / e.sph);
}

// DSS of all tracers of an (element, level) slab plus the extra field;
// at stage 3 the SSP average (Qdp(n0) + 2 Qdp(np1)) / 3 is applied to the
// assembled value (EulerStepFunctorImpl.hpp:379-403).
__global__ __launch_bounds__(256) void euler_dss_kernel(EulerArgs a, const double *in, double *out, int with_fdss,
                                                        int avg) {
  size_t slab;
  int ie, k, pt;
  if (!slab_ie_k((size_t)a.nelem * a.nlev, a.nlev, slab, ie, k, pt)) return;
  const int L = a.nlev, Q = a.qsize;
  const int *src = a.src + ((size_t)ie * 16 + pt) * 3;
  const int s0 = src[0], s1 = src[1], s2 = src[2];
  const double r = a.m.rspheremp[(size_t)ie * 16 + pt];
  const size_t es = (size_t)Q * L * 16;
  for (int q = 0; q < Q; ++q) {
    const size_t off = ((size_t)q * L + k) * 16;
    double x = in[(size_t)ie * es + off + pt];
    if (s0 >= 0) x += in[(size_t)(s0 >> 4) * es + off + (s0 & 15)];
    if (s1 >= 0) x += in[(size_t)(s1 >> 4) * es + off + (s1 & 15)];
    if (s2 >= 0) x += in[(size_t)(s2 >> 4) * es + off + (s2 & 15)];
    x *= r;
    if (avg) x = (a.qdp_avg_n0[(size_t)ie * es + off + pt] + (3.0 - 1) * x) / 3.0;
    out[(size_t)ie * es + off + pt] = x;
  }
  if (with_fdss) {
    const size_t e3 = (size_t)L * 16, off = (size_t)k * 16;
    double x = a.fdss_pre[(size_t)ie * e3 + off + pt];
    if (s0 >= 0) x += a.fdss_pre[(size_t)(s0 >> 4) * e3 + off + (s0 & 15)];
    if (s1 >= 0) x += a.fdss_pre[(size_t)(s1 >> 4) * e3 + off + (s1 & 15)];
    if (s2 >= 0) x += a.fdss_pre[(size_t)(s2 >> 4) * e3 + off + (s2 & 15)];
    x *= r;
    a.fdss[(size_t)ie * e3 + off + pt] = x;
  }
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
    const double addmass = slab_any(delta != 0.0) ? slab_sum_serial(delta * c) : 0.0;
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
  const double addmass = slab_any(delta != 0.0) ? slab_sum_serial(delta * c) : 0.0;
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

// Setup phase (:585-626) + tracer phase (:571-687): per (element, level)
// slab, loop over all tracers.
__global__ __launch_bounds__(256) void euler_advect_kernel(EulerArgs a) {
  size_t slab;
  int ie, k, pt;
  if (!slab_ie_k((size_t)a.nelem * a.nlev, a.nlev, slab, ie, k, pt)) return;
  const int L = a.nlev, Q = a.qsize;
  ElemLane e;
  load_elem(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  const size_t s = slab * 16 + pt, v = slab * 32 + pt;
  // setup
  const double rhsmdt = a.rhs_multiplier * a.dt;
  const double dp = a.derived_dp[s] - rhsmdt * a.divdp_proj[s];
  const double vs0 = a.vn0[v] / dp, vs1 = a.vn0[v + 16] / dp;
  double dpdissk = dp - a.dt * a.divdp[s];
  const bool add_ps_diss = a.nu_p > 0 && a.rhs_viss != 0.0;
  if (add_ps_diss) {
    const double diss_fac = -a.rhs_viss * a.dt * a.nu_q;
    dpdissk += diss_fac * a.dpdiss_bih[s] / e.sph;
  }
  a.fdss_pre[s] = a.fdss[s] * e.sph;
  const bool add_hv = a.rhs_viss != 0.0;
  const double c = e.sph * dpdissk;
  for (int q = 0; q < Q; ++q) {
    const size_t qi = (((size_t)ie * Q + q) * L + k) * 16 + pt;
    const double qd = a.qdp_n0[qi];
    const double q0 = add_hv ? a.qbih[qi] : 0;
    double ptens = sph_div_update(e, d, -a.dt, vs0, vs1, qd, q0);
    if (a.limiter_option == 8 || a.limiter_option == 9) {
      double x = ptens / dpdissk;
      double s0, s1;
      slab_sum_serial2(x * c, c, s0, s1);
      if (!(s1 <= 0)) {
        double *ql = a.qlim + (((size_t)ie * Q + q) * 2) * L;
        double minp = ql[k], maxp = ql[L + k];
        bool wmin = false, wmax = false;
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
        if (pt == 0) {
          if (wmin) ql[k] = minp;
          if (wmax) ql[L + k] = maxp;
        }
        if (a.limiter_option == 8) limit8_slab(s0, minp, maxp, x, c);
        else limit9_slab(minp, maxp, x, c);
        ptens = x * dpdissk;
      }
    }
    a.qdp_pre[qi] = e.sph * ptens;
  }
}

void launch_euler_qminmax(hipStream_t st, const EulerArgs &a, int write_q) {
  const size_t nq = (size_t)a.nelem * a.qsize * a.nlev;
  hipLaunchKernelGGL(euler_qminmax_kernel, dim3((nq + 15) / 16), dim3(256), 0, st, a, write_q);
}

void launch_euler_nbr_minmax(hipStream_t st, const EulerArgs &a) {
  const size_t n = (size_t)a.nelem * a.qsize * a.nlev;
  hipLaunchKernelGGL(euler_nbr_minmax_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_euler_biharmonic(hipStream_t st, const EulerArgs &a) {
  const size_t nq = (size_t)a.nelem * a.qsize * a.nlev;
  hipLaunchKernelGGL(euler_bih_pre_kernel, dim3((nq + 15) / 16), dim3(256), 0, st, a);
  const size_t ns = (size_t)a.nelem * a.nlev;
  hipLaunchKernelGGL(euler_dss_kernel, dim3((ns + 15) / 16), dim3(256), 0, st, a, (const double *)a.qbih_pre, a.qbih,
                     0, 0);
  hipLaunchKernelGGL(euler_bih_post_kernel, dim3((nq + 15) / 16), dim3(256), 0, st, a);
}

void launch_euler_advect(hipStream_t st, const EulerArgs &a) {
  const size_t ns = (size_t)a.nelem * a.nlev;
  hipLaunchKernelGGL(euler_advect_kernel, dim3((ns + 15) / 16), dim3(256), 0, st, a);
}

void launch_euler_dss(hipStream_t st, const EulerArgs &a, int avg) {
  const size_t ns = (size_t)a.nelem * a.nlev;
  hipLaunchKernelGGL(euler_dss_kernel, dim3((ns + 15) / 16), dim3(256), 0, st, a, (const double *)a.qdp_pre,
                     a.qdp_np1, 1, avg);
}

}  // namespace hx
