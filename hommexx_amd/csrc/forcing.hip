// CAM forcing (ftype 0: state and tracers, ftype 2: state only), applied at
// the start of a remap interval (CamForcing.cpp:20-166 in the reference).
#include "kernels.hpp"

namespace hx {

// T(n0) += dt*FT, v(n0) += dt*FM
__global__ void state_forcing_kernel(size_t n, double dt, const double *__restrict__ ft,
                                     const double *__restrict__ fm, double *__restrict__ t, double *__restrict__ v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  t[i] += dt * ft[i];
  const size_t row = i >> 4, pt = i & 15;
  v[row * 32 + pt] += dt * fm[row * 32 + pt];
  v[row * 32 + 16 + pt] += dt * fm[row * 32 + 16 + pt];
}

// moist: ps_v(n0) += sum_k of the clipped dt*FQ of tracer 0 (serial over k)
__global__ void tracer_forcing_ps_kernel(int nelem, int nlev, int qsize, double dt, const double *__restrict__ fq,
                                         const double *__restrict__ qdp, double *__restrict__ ps) {
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (size_t)nelem * 16) return;
  const int ie = (int)(col >> 4), pt = (int)(col & 15);
  const size_t base = (size_t)ie * qsize * nlev * 16 + pt;  // tracer 0
  double acc = 0.0;
  for (int k = 0; k < nlev; ++k) {
    double v1 = dt * fq[base + (size_t)k * 16];
    const double qdp_s = qdp[base + (size_t)k * 16];
    if (qdp_s + v1 < 0.0 && v1 < 0.0) v1 = qdp_s < 0.0 ? 0.0 : -qdp_s;
    acc += v1;
  }
  ps[col] += acc;
}

// Qdp += clipped dt*FQ; Q = Qdp/dp(ps_v)
__global__ void tracer_forcing_kernel(int nelem, int nlev, int qsize, double dt, const double *__restrict__ fq,
                                      double *__restrict__ qdp, double *__restrict__ Q,
                                      const double *__restrict__ hyai, const double *__restrict__ hybi, double ps0,
                                      const double *__restrict__ ps) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nelem * qsize * nlev * 16) return;
  const int pt = (int)(i & 15);
  const int k = (int)((i >> 4) % nlev);
  const int ie = (int)((i >> 4) / ((size_t)nlev * qsize));
  double v1 = dt * fq[i];
  double q = qdp[i];
  if (q + v1 < 0.0 && v1 < 0.0) v1 = q < 0.0 ? 0.0 : -q;
  q += v1;
  qdp[i] = q;
  const double dp = (hyai[k + 1] - hyai[k]) * ps0 + (hybi[k + 1] - hybi[k]) * ps[(size_t)ie * 16 + pt];
  Q[i] = q / dp;
}

void launch_state_forcing(hipStream_t st, int nelem, int nlev, double dt, const double *ft, const double *fm,
                          double *t, double *v) {
  const size_t n = (size_t)nelem * nlev * 16;
  hipLaunchKernelGGL(state_forcing_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, dt, ft, fm, t, v);
}

void launch_tracer_forcing(hipStream_t st, int nelem, int nlev, int qsize, bool moist, double dt, const double *fq,
                           double *qdp, double *Q, const double *hyai, const double *hybi, double ps0, double *ps) {
  if (moist && qsize > 0) {
    const size_t nc = (size_t)nelem * 16;
    hipLaunchKernelGGL(tracer_forcing_ps_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, nelem, nlev, qsize, dt, fq,
                       (const double *)qdp, ps);
  }
  const size_t n = (size_t)nelem * qsize * nlev * 16;
  if (n)
    hipLaunchKernelGGL(tracer_forcing_kernel, dim3((n + 255) / 256), dim3(256), 0, st, nelem, nlev, qsize, dt, fq, qdp,
                       Q, hyai, hybi, ps0, (const double *)ps);
}

}  // namespace hx
