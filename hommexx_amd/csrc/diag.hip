// Diagnostics of the reference's Diagnostics class (Diagnostics.cpp:37-185):
// per-column level sums that the Fortran driver turns into its energy and
// mass reports.  One lane per column (and tracer); levels are summed in
// order, as the reference's host loops do.
#include "kernels.hpp"

namespace hx {

constexpr double CPWATER_VAPOR = 1870.0;

// prim_energy_halftimes (Diagnostics.cpp:90-185) at dynamics level (v, t, ps)
// and tracer level qdp (tracer 0 gives cp* with use_cpstar):
// out[0..3][col] = sum cp* T dp, sum (cp* - cp) T dp, sum (u^2 + v^2)/2 dp,
// sum phis dp, with dp = (hyai(k+1) - hyai(k)) ps0 + (hybi(k+1) - hybi(k)) ps.
__global__ __launch_bounds__(256) void diag_energy_kernel(DiagArgs a) {
  const size_t n = (size_t)a.nelem * 16;
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  const int L = a.nlev;
  const size_t ie = col >> 4, pt = col & 15;
  const double ps = a.ps[col], phis = a.phis[col];
  const double *u = a.v + ie * L * 32 + pt, *w = u + 16, *t = a.t + ie * L * 16 + pt;
  const double *q0 = a.use_cpstar ? a.qdp + ie * a.qsize * L * 16 + pt : nullptr;
  double ien = 0.0, iwet = 0.0, ken = 0.0, pen = 0.0;
  for (int k = 0; k < L; ++k) {
    const double dpt1 = (a.hyai[k + 1] - a.hyai[k]) * a.ps0 + (a.hybi[k + 1] - a.hybi[k]) * ps;
    double cps = CP;
    if (a.use_cpstar) {
      const double qv = q0[(size_t)k * 16] / dpt1;
      cps = CP * (1.0 + (CPWATER_VAPOR / CP - 1.0) * qv);
    }
    const double tk = t[(size_t)k * 16], uk = u[(size_t)k * 32], wk = w[(size_t)k * 32];
    ien += cps * tk * dpt1;
    iwet += (cps - CP) * tk * dpt1;
    ken += (uk * uk + wk * wk) * 0.5 * dpt1;  // std::pow(x, 2): the correctly rounded square
    pen += phis * dpt1;
  }
  a.out[col] = ien;
  a.out[n + col] = iwet;
  a.out[2 * n + col] = ken;
  a.out[3 * n + col] = pen;
}

// prim_diag_scalars (Diagnostics.cpp:37-88): per column and tracer,
// out[0][i] = sum Qdp Q, out[1][i] = sum Qdp, i = (ie * qsize + q) * 16 + pt.
__global__ __launch_bounds__(256) void diag_scalars_kernel(DiagArgs a) {
  const size_t n = (size_t)a.nelem * a.qsize * 16;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int L = a.nlev;
  const size_t base = (i >> 4) * L * 16 + (i & 15);
  const double *qdp = a.qdp + base, *Q = a.Q + base;
  double sqq = 0.0, sq = 0.0;
  for (int k = 0; k < L; ++k) {
    const double x = qdp[(size_t)k * 16];
    sqq += x * Q[(size_t)k * 16];
    sq += x;
  }
  a.out[i] = sqq;
  a.out[n + i] = sq;
}

void launch_diag_energy(hipStream_t st, const DiagArgs &a) {
  const size_t n = (size_t)a.nelem * 16;
  hipLaunchKernelGGL(diag_energy_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_diag_scalars(hipStream_t st, const DiagArgs &a) {
  const size_t n = (size_t)a.nelem * a.qsize * 16;
  if (n) hipLaunchKernelGGL(diag_scalars_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace hx
