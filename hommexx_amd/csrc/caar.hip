// Dynamics right-hand side and update for one RK stage (CAAR), plus the
// small element-wise kernels of the step (dp3d from ps, prim_step init,
// the (5u1-u0)/4 stage combination).
//
// CAAR: a block holds 4 elements x 16 GLL columns x CNC level chunks; the
// 16 lanes of one element and chunk form a slab group, so the horizontal
// derivatives at each level are cross-lane operations inside it.  Each chunk
// walks its own levels; the column scans stay serial recurrences in the
// reference's order (CaarFunctorImpl.hpp:621-650, 689-729, 854-889): a chunk
// first re-runs the recurrence over the levels above it (pressure), or reads
// the other chunks' terms after a block barrier (omega, geopotential).
//   A: p; vdp, vn0 += w vdp, div(vdp) -> divv scratch
//   B: omega/p (serial sum of divv above), T(np1), dp3d(np1),
//      Rgas*T_v*dp/(2p) -> phi
//   C: phi = phis + 2*sum_below + own term (serial from the bottom)
//   D: energy gradient, vorticity, v(np1)
// With rsplit == 0 phases B and D also carry the Eulerian vertical advection
// (eta_dot_dpdn from the column sums of div(v dp)).
// XL (rsplit > 0, nlev <= CAAR_LDS_MAX_NLEV): div(v dp) and then the phi
// terms live in one LDS array of the block's 4 columns x nlev, never in HBM
// (phi is internal to the stage); otherwise both go through global scratch.
// Outputs go to pre-exchange buffers; the DSS kernel writes the state.
#include "kernels.hpp"

namespace hx {

constexpr int CNC = 4;  // level chunks per column
// the LDS array of 4 columns x nlev levels fits the 64 KB a block may ask for
constexpr int CAAR_LDS_MAX_NLEV = 128;

__device__ __forceinline__ double virtual_t(const CaarArgs &a, const double *qv, int k, double t, double dp) {
  if (!qv) return t;
  double qt = qv[(size_t)k * 16] / dp;
  qt *= (RWATER_VAPOR / RGAS - 1.0);
  qt += 1.0;
  return t * qt;
}

template <bool RS0, bool XL>  // rsplit == 0: Eulerian vertical advection; XL: divv/phi in LDS
__global__ __launch_bounds__(64 * CNC) void caar_kernel(CaarArgs a) {
  extern __shared__ double xs[];
  const int pt = threadIdx.x & 15;
  const int chunk = threadIdx.x >> 6;
  const int ie = blockIdx.x * 4 + ((threadIdx.x >> 4) & 3);
  const bool on = ie < a.nelem;
  const int L = a.nlev;
  const int cs = (L + CNC - 1) / CNC;
  const int k0 = chunk * cs < L ? chunk * cs : L;
  const int k1 = k0 + cs < L ? k0 + cs : L;
  ElemLane e;
  DvvLane d;
  if (on) {
    load_elem(a.m, ie, pt, e);
    load_dvv(a.m.dvv, pt, d);
  }
  const size_t sb = (size_t)(on ? ie : 0) * L * 16 + pt;  // scalar [ie][lev][pt]
  const size_t vb = (size_t)(on ? ie : 0) * L * 32 + pt;  // vector [ie][lev][c][pt]
  // this column's div(v dp) / phi levels, stride 16
  double *const divv = XL ? xs + (size_t)((threadIdx.x >> 4) & 3) * L * 16 + pt : a.divv + sb;
  double *const phi = XL ? divv : a.phi + sb;
  const double eta = a.eta_ave_w, dt = a.dt;
  const double *qv = a.qdp_n0 ? a.qdp_n0 + (size_t)(on ? ie : 0) * a.qsize * L * 16 + pt : nullptr;

  // pressure at the top of the chunk: the serial recurrence over the levels above
  double p_top = a.p_top, dp_top = 0.0;
  if (on)
#pragma unroll 8
    for (int k = 0; k < k0; ++k) {
      const double dp = a.dp_n0[sb + (size_t)k * 16];
      p_top = p_top + 0.5 * (dp_top + dp);
      dp_top = dp;
    }

  // ---------------------------------------------------------------- A
  // Phases A, B and D load level k+1's inputs while level k is computed (the
  // loads of y are issued first; x = y at the end waits for them)
  auto kn = [&](int k) { return k + 1 < k1 ? k + 1 : k; };
  if (on) {
    double x[5], y[5];  // dp, u, w, vn0 (2)
    auto ldA = [&](int k, double (&r)[5]) {
      const size_t s = sb + (size_t)k * 16, v = vb + (size_t)k * 32;
      r[0] = a.dp_n0[s];
      r[1] = a.v_n0[v];
      r[2] = a.v_n0[v + 16];
      r[3] = a.vn0[v];
      r[4] = a.vn0[v + 16];
    };
    if (k0 < k1) ldA(k0, x);
    for (int k = k0; k < k1; ++k) {
      ldA(kn(k), y);
      const size_t v = vb + (size_t)k * 32;
      const double dp = x[0];
      const double u = x[1], w = x[2];
      const double vdp0 = u * dp, vdp1 = w * dp;
      a.vn0[v] = x[3] + eta * vdp0;
      a.vn0[v + 16] = x[4] + eta * vdp1;
      divv[(size_t)k * 16] = sph_div(e, d, vdp0, vdp1);
#pragma unroll
      for (int i = 0; i < 5; ++i) x[i] = y[i];
    }
  }
  __syncthreads();
  // ---------------------------------------------------------------- B
  // rsplit == 0: eta_dot_dpdn(k) = hybi(k) * S(L-1) - S(k-1), eta(0) = 0,
  // with S the serial column sums of div(v dp) (:264-301)
  double stotal = 0.0;
  if (on && RS0)
    for (int k = 0; k < L; ++k) stotal += divv[(size_t)k * 16];
  auto eta_of = [&](int k, double s_below) {  // s_below = S(k-1)
    return k == 0 ? 0.0 : (k < L ? a.hybi[k] * stotal - s_below : 0.0);
  };
  // vertical advection of x at level k (preq_vertadv, :497-597)
  auto vadv = [&](const double *x, size_t stride, int k, double dpk, double ek, double ek1) {
    const double xk = x[(size_t)k * stride];
    if (k == 0) {
      const double facp = (0.5 * 1 / dpk) * ek1;
      return facp * (x[(size_t)(k + 1) * stride] - xk);
    }
    if (k == L - 1) {
      const double facm = (0.5 * (1 / dpk)) * ek;
      return facm * (xk - x[(size_t)(k - 1) * stride]);
    }
    const double facp = 0.5 * (1 / dpk) * ek1;
    const double facm = 0.5 * (1 / dpk) * ek;
    return facp * (x[(size_t)(k + 1) * stride] - xk) + facm * (xk - x[(size_t)(k - 1) * stride]);
  };
  double integ = 0.0;
  if (on)
    for (int k = 0; k < k0; ++k) integ = integ + divv[(size_t)k * 16];
  // XL: the phi terms below overwrite div(v dp) level by level, so every
  // chunk first reads the levels above it
  if (XL) __syncthreads();
  if (on) {
    double p_prev = p_top, dp_prev = dp_top;
    double x[7], y[7];  // dp, u, w, T, omega, T(nm1), dp3d(nm1)
    auto ldB = [&](int k, double (&r)[7]) {
      const size_t s = sb + (size_t)k * 16, v = vb + (size_t)k * 32;
      r[0] = a.dp_n0[s];
      r[1] = a.v_n0[v];
      r[2] = a.v_n0[v + 16];
      r[3] = a.t_n0[s];
      r[4] = a.omega[s];
      r[5] = a.t_nm1[s];
      r[6] = a.dp_nm1[s];
    };
    if (k0 < k1) ldB(k0, x);
    for (int k = k0; k < k1; ++k) {
      ldB(kn(k), y);
      const size_t s = sb + (size_t)k * 16;
      const double dp = x[0];
      const double u = x[1], w = x[2];
      const double t = x[3];
      const double tv = virtual_t(a, qv, k, t, dp);
      const double p = p_prev + 0.5 * (dp_prev + dp);
      p_prev = p;
      dp_prev = dp;
      const double dvk = divv[(size_t)k * 16];
      double ek = 0.0, ek1 = 0.0, tva = 0.0;
      if (RS0) {
        ek = eta_of(k, integ);
        ek1 = eta_of(k + 1, integ + dvk);
        tva = vadv(a.t_n0 + sb, 16, k, dp, ek, ek1);
        a.eta_acc[s] += eta * ek;
      }
      double gp0, gp1;
      sph_grad(e, d, p, gp0, gp1);
      const double vgrad_p = u * gp0 + w * gp1;
      const double om = (vgrad_p - (integ + 0.5 * dvk)) / p;
      integ = integ + dvk;
      a.omega[s] = x[4] + eta * om;
      double gt0, gt1;
      sph_grad(e, d, t, gt0, gt1);
      const double vgrad_t = u * gt0 + w * gt1;
      const double ttens = (RS0 ? -tva : 0) - vgrad_t + KAPPA * tv * om;
      double temp = ttens * dt + x[5];
      temp *= e.sph;
      a.out_t[s] = temp;
      double tmp = ek1;
      tmp += dvk;
      tmp -= ek;
      tmp = x[6] - tmp * dt;
      a.out_dp[s] = e.sph * tmp;
      phi[(size_t)k * 16] = RGAS * tv * (dp * 0.5 / p);
#pragma unroll
      for (int i = 0; i < 7; ++i) x[i] = y[i];
    }
  }
  __syncthreads();
  // ---------------------------------------------------------------- C
  double integration = 0.0;
  if (on)
    for (int k = L - 1; k >= k1; --k) integration = integration + phi[(size_t)k * 16];
  __syncthreads();  // every chunk has read the terms below it before they are overwritten
  if (on)
    for (int k = k1 - 1; k >= k0; --k) {
      const double r = phi[(size_t)k * 16];
      phi[(size_t)k * 16] = e.phis + 2.0 * integration + r;
      integration = integration + r;
    }
  // ---------------------------------------------------------------- D
  if (on) {
    double p_prev = p_top, dp_prev = dp_top;
    double integ = 0.0;
    if (RS0)
      for (int k = 0; k < k0; ++k) integ = integ + divv[(size_t)k * 16];
    double x[6], y[6];  // dp, u, w, T, v(nm1) (2)
    auto ldD = [&](int k, double (&r)[6]) {
      const size_t s = sb + (size_t)k * 16, v = vb + (size_t)k * 32;
      r[0] = a.dp_n0[s];
      r[1] = a.v_n0[v];
      r[2] = a.v_n0[v + 16];
      r[3] = a.t_n0[s];
      r[4] = a.v_nm1[v];
      r[5] = a.v_nm1[v + 16];
    };
    if (k0 < k1) ldD(k0, x);
    for (int k = k0; k < k1; ++k) {
      ldD(kn(k), y);
      const size_t v = vb + (size_t)k * 32;
      const double dp = x[0];
      const double u = x[1], w = x[2];
      const double t = x[3];
      const double tv = virtual_t(a, qv, k, t, dp);
      const double p = p_prev + 0.5 * (dp_prev + dp);
      p_prev = p;
      dp_prev = dp;
      double gp0, gp1;
      sph_grad(e, d, p, gp0, gp1);
      double eg0 = RGAS * (tv / p) * gp0;
      double eg1 = RGAS * (tv / p) * gp1;
      const double ephi = 0.5 * (u * u + w * w) + phi[(size_t)k * 16];
      sph_grad_update(e, d, ephi, eg0, eg1);
      double vort = sph_vort(e, d, u, w);
      vort += e.fcor;
      double va0 = 0.0, va1 = 0.0;
      if (RS0) {
        const double dv = divv[(size_t)k * 16];
        const double ek = eta_of(k, integ), ek1 = eta_of(k + 1, integ + dv);
        integ = integ + dv;
        va0 = vadv(a.v_n0 + vb, 32, k, dp, ek, ek1);
        va1 = vadv(a.v_n0 + vb + 16, 32, k, dp, ek, ek1);
      }
      eg0 *= -1;
      eg0 += (RS0 ? -va0 : 0) + w * vort;
      eg1 *= -1;
      eg1 += (RS0 ? -va1 : 0) - u * vort;
      eg0 *= dt;
      eg0 += x[4];
      eg1 *= dt;
      eg1 += x[5];
      a.out_v[v] = e.sph * eg0;
      a.out_v[v + 16] = e.sph * eg1;
#pragma unroll
      for (int i = 0; i < 6; ++i) x[i] = y[i];
    }
  }
}

void launch_caar(hipStream_t st, const CaarArgs &a) {
  const int blocks = (a.nelem + 3) / 4;
  if (a.rsplit0) hipLaunchKernelGGL((caar_kernel<true, false>), dim3(blocks), dim3(64 * CNC), 0, st, a);
  else if (a.nlev <= CAAR_LDS_MAX_NLEV)
    hipLaunchKernelGGL((caar_kernel<false, true>), dim3(blocks), dim3(64 * CNC), (size_t)4 * a.nlev * 16 * sizeof(double),
                       st, a);
  else hipLaunchKernelGGL((caar_kernel<false, false>), dim3(blocks), dim3(64 * CNC), 0, st, a);
}

// ------------------------------------------------------------------ small
__global__ void dp3d_from_ps_kernel(int nelem, int nlev, const double *__restrict__ hyai,
                                    const double *__restrict__ hybi, double ps0, const double *__restrict__ ps,
                                    double *__restrict__ dp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)nelem * nlev * 16;
  if (i >= n) return;
  const int pt = i & 15;
  const int k = (i >> 4) % nlev;
  const int ie = (int)((i >> 4) / nlev);
  const double da = hyai[k + 1] - hyai[k], db = hybi[k + 1] - hybi[k];
  dp[i] = da * ps0 + db * ps[(size_t)ie * 16 + pt];
}

void launch_dp3d_from_ps(hipStream_t st, int nelem, int nlev, const double *hyai, const double *hybi, double ps0,
                         const double *ps, double *dp) {
  const size_t n = (size_t)nelem * nlev * 16;
  hipLaunchKernelGGL(dp3d_from_ps_kernel, dim3((n + 255) / 256), dim3(256), 0, st, nelem, nlev, hyai, hybi, ps0, ps,
                     dp);
}

__global__ void prim_step_init_kernel(size_t n, bool zero_diss, const double *__restrict__ dp_n0,
                                      double *__restrict__ eta_dot, double *__restrict__ vn0,
                                      double *__restrict__ omega, double *__restrict__ dpa, double *__restrict__ dpb,
                                      double *__restrict__ derived_dp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  eta_dot[i] = 0;
  omega[i] = 0;
  const size_t row = i >> 4, pt = i & 15;
  vn0[row * 32 + pt] = 0;
  vn0[row * 32 + 16 + pt] = 0;
  if (zero_diss) {
    dpa[i] = 0;
    dpb[i] = 0;
  }
  derived_dp[i] = dp_n0[i];
}

void launch_prim_step_init(hipStream_t st, int nelem, int nlev, bool zero_diss, const double *dp_n0,
                           double *eta_dot, double *vn0, double *omega, double *dpa, double *dpb,
                           double *derived_dp) {
  const size_t n = (size_t)nelem * nlev * 16;
  hipLaunchKernelGGL(prim_step_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, zero_diss, dp_n0, eta_dot,
                     vn0, omega, dpa, dpb, derived_dp);
}

// nm1 <- (5*nm1 - n0)/4 for T, v, dp3d (prim_advance_exp.cpp:139-156)
__global__ void rk_combo_kernel(size_t n, double *__restrict__ v_nm1, double *__restrict__ t_nm1,
                                double *__restrict__ dp_nm1, const double *__restrict__ v_n0,
                                const double *__restrict__ t_n0, const double *__restrict__ dp_n0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  t_nm1[i] = (5.0 * t_nm1[i] - t_n0[i]) / 4.0;
  const size_t row = i >> 4, pt = i & 15;
  const size_t v0 = row * 32 + pt, v1 = v0 + 16;
  v_nm1[v0] = (5.0 * v_nm1[v0] - v_n0[v0]) / 4.0;
  v_nm1[v1] = (5.0 * v_nm1[v1] - v_n0[v1]) / 4.0;
  dp_nm1[i] = (5.0 * dp_nm1[i] - dp_n0[i]) / 4.0;
}

void launch_rk_combo(hipStream_t st, int nelem, int nlev, double *v_nm1, double *t_nm1, double *dp_nm1,
                     const double *v_n0, const double *t_n0, const double *dp_n0) {
  const size_t n = (size_t)nelem * nlev * 16;
  hipLaunchKernelGGL(rk_combo_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, v_nm1, t_nm1, dp_nm1, v_n0, t_n0,
                     dp_n0);
}

}  // namespace hx
