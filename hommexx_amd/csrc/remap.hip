// Vertical remap of (v dp, T dp, Qdp) from the Lagrangian levels back to the
// reference hybrid levels with monotone PPM (RemapFunctor.hpp:226-331,
// PpmRemap.hpp:78-673 in the reference), plus update_q (prim_driver.cpp:171-206).
//
// The reference builds per-column arrays (pio, pin, ppmdx[10], kid, z2, ao,
// dma, ai, parabola coefficients) and then remaps; here one top-to-bottom
// sweep produces the same numbers with O(1) state: source cell j's parabola
// needs ai(j-1), ai(j), which need dma(j-1..j+1) and a, dx up to j+2, so a
// four-cell window suffices, and the target interface k+1 lies in source cell
// kid[k] >= k-1 (kid is found by the same forward search).  Every value is
// computed with the reference's expression and operand order; grid factors
// that depend on the thicknesses only (the reference's ppmdx rows) are
// evaluated once per column and shared by a batch of fields.
#include "kernels.hpp"

namespace hx {

// ps_v = sum(dp3d) + hyai0*ps0 and the thickness check (ComputeThicknessTag).
// rsplit == 0: the source thickness of the tracer remap is the target
// thickness plus dt*(eta_dot_dpdn(k+1) - eta_dot_dpdn(k)) (RemapFunctor.hpp:62-90).
__global__ __launch_bounds__(256) void remap_thickness_kernel(RemapArgs a) {
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (size_t)a.nelem * 16) return;
  const int ie = (int)(col >> 4), pt = (int)(col & 15);
  const double *dp = a.dp + (size_t)ie * a.nlev * 16 + pt;
  double s = 0.0;
  bool bad = false;
  for (int k = 0; k < a.nlev; ++k) {
    const double x = dp[(size_t)k * 16];
    s += x;
    if (a.nstates) bad |= isnan(x) || x < 0.0;
  }
  s += a.hyai0 * a.ps0;
  a.ps[col] = s;
  if (!a.nstates) {
    const double *eta = a.eta + (size_t)ie * a.nlev * 16 + pt;
    double *src = a.src + (size_t)ie * a.nlev * 16 + pt;
    for (int k = 0; k < a.nlev; ++k) {
      const double tgt = (a.hyai[k + 1] - a.hyai[k]) * a.ps0 + (a.hybi[k + 1] - a.hybi[k]) * s;
      const double en = k + 1 < a.nlev ? eta[(size_t)(k + 1) * 16] : 0;
      const double delta = en - eta[(size_t)k * 16];
      const double x = tgt + a.dt * delta;
      src[(size_t)k * 16] = x;
      bad |= isnan(x) || x < 0.0;
    }
  }
  if (bad) atomicOr(a.error, 1);
}

__device__ __forceinline__ double dmin3(double a, double b, double c) {
  const double bc = b < c ? b : c;
  return a < bc ? a : bc;
}

__device__ __forceinline__ double integrate_parabola(double sq, double lin, double cst, double x1, double x2) {
  return (cst * (x2 - x1) + lin * (x2 * x2 - x1 * x1) / 2.0) + sq * (x2 * x2 * x2 - x1 * x1 * x1) / 3.0;
}

// dma(j) from a(j-1..j+1), dx(j-1..j+1)
__device__ __forceinline__ double ppm_dma(double am1, double a0, double a1, double xm1, double x0, double x1) {
  if ((a1 - a0) * (a0 - am1) > 0.0) {
    const double g0 = x0 / (xm1 + x0 + x1);
    const double g1 = (2.0 * xm1 + x0) / (x0 + x1);
    const double g2 = (x0 + 2.0 * x1) / (xm1 + x0);
    const double da = g0 * (g1 * (a1 - a0) + g2 * (a0 - am1));
    return dmin3(fabs(da), 2.0 * fabs(a0 - am1), 2.0 * fabs(a1 - a0)) * copysign(1.0, da);
  }
  return 0.0;
}

// the same with the grid factors g0, g1, g2 precomputed
__device__ __forceinline__ double ppm_dma_g(double am1, double a0, double a1, double g0, double g1, double g2) {
  if ((a1 - a0) * (a0 - am1) > 0.0) {
    const double da = g0 * (g1 * (a1 - a0) + g2 * (a0 - am1));
    return dmin3(fabs(da), 2.0 * fabs(a0 - am1), 2.0 * fabs(a1 - a0)) * copysign(1.0, da);
  }
  return 0.0;
}

// ai(j) from a(j), a(j+1), dma(j), dma(j+1) and the grid factors of
// dx(j-1..j+2): g3, g4, g56 = g5*(g6-g7), g8, g9 (the product associates
// left to right, so g5*(g6-g7)*(a1-a0) == g56*(a1-a0) exactly)
__device__ __forceinline__ double ppm_ai_g(double a0, double a1, double dma0, double dma1, const double *g) {
  return a0 + g[0] * (a1 - a0) + g[1] * (g[2] * (a1 - a0) - g[3] * dma1 + g[4] * dma0);
}

__device__ __forceinline__ void ppm_parabola(double al, double ar, double a, double &c0, double &c1, double &c2) {
  if ((ar - a) * (a - al) <= 0.) {
    al = a;
    ar = a;
  }
  // the two monotonicity tests share (ar - al) and (ar - al)^2 / 6 unless
  // the first one moves al (same operations and rounding as evaluating each
  // test in full)
  double d = ar - al, q = d * d / 6.0;
  if (d * (a - (al + ar) / 2.0) > q) {
    al = 3.0 * a - 2.0 * ar;
    d = ar - al;
    q = d * d / 6.0;
  }
  if (d * (a - (al + ar) / 2.0) < -q) ar = 3.0 * a - 2.0 * al;
  c0 = 1.5 * a - (al + ar) / 4.0;
  c1 = ar - al;
  c2 = 3.0 * (-2.0 * a + (al + ar));
}

// One lane = one GLL column and a batch of RFB fields (field 0,1: v
// components, 2: T, >= 3: tracer field-3).  A wave is 64 columns (four
// elements) of one field batch, so the field kinds are wave-uniform.  The
// grid factors, the target search (kid, z2) and the control flow of the sweep
// depend on the column only and are shared by the batch.
template <int RFB>  // fields per lane
__global__ __launch_bounds__(256) void remap_kernel(RemapArgs a, int nfb) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int w = lb * 4 + (threadIdx.x >> 6);  // global wave
  const int eg = w / nfb, fb = w % nfb;
  const int ie = eg * 4 + ((threadIdx.x >> 4) & 3);
  if (ie >= a.nelem) return;
  const int pt = threadIdx.x & 15;
  const int nf = 3 + a.qsize;  // in the field numbering below (0,1: v, 2: T, 3..: tracers)
  const int f0 = 3 - a.nstates;  // rsplit == 0: tracers only
  const size_t col = (size_t)ie * 16 + pt;
  const int L = a.nlev;
  const double *src = (a.nstates ? a.dp : a.src) + (size_t)ie * L * 16 + pt;
  double *var[RFB];
  double *Qout[RFB];
  size_t vs[RFB];
  bool on[RFB], state[RFB];
#pragma unroll
  for (int i = 0; i < RFB; ++i) {
    const int f = f0 + fb * RFB + i;
    on[i] = f < nf;
    state[i] = f < 3;
    vs[i] = f < 2 ? 32 : 16;
    Qout[i] = nullptr;
    if (f < 2) var[i] = a.v + (size_t)ie * L * 32 + f * 16 + pt;
    else if (f == 2) var[i] = a.t + (size_t)ie * L * 16 + pt;
    else if (f < nf) {
      var[i] = a.qdp + (((size_t)ie * a.qsize + (f - 3)) * L) * 16 + pt;
      if (a.Q) Qout[i] = a.Q + (((size_t)ie * a.qsize + (f - 3)) * L) * 16 + pt;
    } else var[i] = nullptr;
  }
  const double ps = a.ps[col];
  const int alg = a.remap_alg;
  auto tgt_of = [&](int k) {
    return (a.hyai[k + 1] - a.hyai[k]) * a.ps0 + (a.hybi[k + 1] - a.hybi[k]) * ps;
  };
  // pio(L): serial sum of the source thicknesses
  double pio_L = 0.0;
  for (int c = 0; c < L; ++c) pio_L += src[(size_t)c * 16];

  // ------------------------------------------------------------ window
  // Fortran-style cell indices j = 1..L, mirrored (or constant) ghosts.
  // (Xm1, X0, X1, X2) = dx(j-1..j+2); per field (Am1..A2) = a(j-1..j+2),
  // V0..V2 = var(j..j+2)
  const double s0 = src[0], s1 = src[16];
  double Xm1 = s1, X0 = s0, X1 = s0, X2 = s1;
  double Am1[RFB], A0[RFB], A1[RFB], A2[RFB], V0[RFB], V1[RFB], V2[RFB];
  double dma_j[RFB], ai_prev[RFB], var_prefix[RFB], mass_prev[RFB];
  double cc0[RFB], cc1[RFB], cc2[RFB];
  double g0, g1, g2;  // dma factors of the current window
  {
    g0 = X0 / (Xm1 + X0 + X1);
    g1 = (2.0 * Xm1 + X0) / (X0 + X1);
    g2 = (X0 + 2.0 * X1) / (Xm1 + X0);
  }
#pragma unroll
  for (int i = 0; i < RFB; ++i) {
    if (!on[i]) continue;
    const double x0 = var[i][0], x1 = var[i][vs[i]];
    const double v0 = state[i] ? x0 * s0 : x0, v1 = state[i] ? x1 * s1 : x1;
    const double a1 = v0 / s0, a2 = v1 / s1;
    Am1[i] = alg == 3 ? a1 : a2;
    A0[i] = a1;
    A1[i] = a1;
    A2[i] = a2;
    V0[i] = 0.0;
    V1[i] = v0;
    V2[i] = v1;
    dma_j[i] = ppm_dma_g(Am1[i], A0[i], A1[i], g0, g1, g2);  // dma(0)
    ai_prev[i] = 0.0;
    var_prefix[i] = 0.0;
    mass_prev[i] = 0.0;
  }

  // ------------------------------------------------------------ targets
  int k = 0;     // next target level
  int kid = -2;  // kid of target k (-2: not yet computed)
  double z2 = 0.0, pin_acc = 0.0;
  int si = 0;
  double pio_si = 0.0, pio_m1 = 0.0, pio_m2 = 0.0;  // pio(si), pio(si-1), pio(si-2)

  auto advance_si = [&]() {
    pio_m2 = pio_m1;
    pio_m1 = pio_si;
    pio_si = si < L ? pio_si + src[(size_t)si * 16] : pio_si + 1.0;  // pio(L+1) = pio(L) + 1
    ++si;
  };
  auto compute_kid = [&]() {
    double pin_k1;
    if (k + 1 < L) {
      pin_acc += tgt_of(k);
      pin_k1 = pin_acc;
    } else {
      pin_k1 = pio_L;  // pin(L) = pio(L)
    }
    while (si < k) advance_si();  // the search for target k starts at interface k
    while (si <= L && pio_si <= pin_k1) advance_si();
    int kk = si;
    double plo = pio_m1, phi = pio_si;  // pio(kk-1), pio(kk)
    if (kk == L + 1) {
      kk = L;
      plo = pio_m2;
      phi = pio_m1;
    }
    kid = kk - 1;
    z2 = kid >= 0 ? (pin_k1 - (plo + phi) * 0.5) / src[(size_t)kid * 16] : 0.0;
  };
  // finish target k, whose upper interface lies in the current cell
  auto finish = [&](double dx) {
    const double tk = tgt_of(k);
#pragma unroll
    for (int i = 0; i < RFB; ++i) {
      if (!on[i]) continue;
      const double integral = integrate_parabola(cc2[i], cc1[i], cc0[i], -0.5, z2);
      const double mass = var_prefix[i] + integral * dx;
      const double out = k > 0 ? mass - mass_prev[i] : mass;
      mass_prev[i] = mass;
      if (state[i]) {
        var[i][(size_t)k * vs[i]] = out / tk;
      } else {
        var[i][(size_t)k * vs[i]] = out;
        if (Qout[i]) Qout[i][(size_t)k * 16] = out / tk;
      }
    }
    ++k;
    kid = -2;
  };

  // ------------------------------------------------------------ sweep
  // cell j+3 enters the window at iteration j; its dx and field values are
  // loaded two iterations ahead so the load latency hides behind sweep steps
  double px = 0.0, pv[RFB], qx = 0.0, qv[RFB];
  auto fetch = [&](int m, double &x, double *v) {
    if (m <= L) {
      x = src[(size_t)(m - 1) * 16];
#pragma unroll
      for (int i = 0; i < RFB; ++i)
        if (on[i]) v[i] = var[i][(size_t)(m - 1) * vs[i]];
    }
  };
  fetch(3, px, pv);
  fetch(4, qx, qv);
  for (int j = 0; j <= L; ++j) {
    double nx = 0.0, nv[RFB];
    fetch(j + 5, nx, nv);
    // dma(j+1) factors: cells j..j+2; ai(j) factors: cells j-1..j+2
    const double h0 = X1 / (X0 + X1 + X2);
    const double h1 = (2.0 * X0 + X1) / (X1 + X2);
    const double h2 = (X1 + 2.0 * X2) / (X0 + X1);
    double gai[5];
    gai[0] = X0 / (X0 + X1);
    gai[1] = 1.0 / (Xm1 + X0 + X1 + X2);
    gai[2] = (2.0 * X0 * X1) / (X0 + X1) * ((Xm1 + X0) / (2.0 * X0 + X1) - (X2 + X1) / (2.0 * X1 + X0));
    gai[3] = X0 * (Xm1 + X0) / (2.0 * X0 + X1);
    gai[4] = X1 * (X1 + X2) / (X0 + 2.0 * X1);
    double dma_j1[RFB], ai_j[RFB];
#pragma unroll
    for (int i = 0; i < RFB; ++i) {
      if (!on[i]) continue;
      dma_j1[i] = ppm_dma_g(A0[i], A1[i], A2[i], h0, h1, h2);            // dma(j+1)
      ai_j[i] = ppm_ai_g(A0[i], A1[i], dma_j[i], dma_j1[i], gai);       // ai(j)
    }
    if (j >= 1) {
      const int c = j - 1;  // physical source cell, Fortran cell j
#pragma unroll
      for (int i = 0; i < RFB; ++i) {
        if (!on[i]) continue;
        ppm_parabola(ai_prev[i], ai_j[i], A0[i], cc0[i], cc1[i], cc2[i]);
        if (alg == 2 && (c <= 1 || c >= L - 2)) {
          cc0[i] = A0[i];
          cc1[i] = 0.0;
          cc2[i] = 0.0;
        }
      }
      // kid is non-decreasing in k (each search continues from the last
      // interface reached, starting no lower than k), so every target is
      // finished while the sweep is at its own cell kid(k)
      while (k < L) {
        if (kid == -2) compute_kid();
        if (kid == c) {
          finish(X0);
        } else if (kid < c) {
          atomicOr(a.error, 2);  // deformation beyond one layer: not supported
          return;
        } else {
          break;
        }
      }
#pragma unroll
      for (int i = 0; i < RFB; ++i)
        if (on[i]) var_prefix[i] += V0[i];
    }
    // shift: bring in cell j+3
    const int m = j + 3;
    double xn;
    if (m <= L) xn = px;
    else if (m == L + 1) xn = X2;  // dx(L+1) = dx(L)
    else xn = X0;                  // dx(L+2) = dx(L-1)
#pragma unroll
    for (int i = 0; i < RFB; ++i) {
      if (!on[i]) continue;
      double an, vn = 0.0;
      if (m <= L) {
        const double x = pv[i];
        vn = state[i] ? x * xn : x;
        an = vn / xn;
      } else if (m == L + 1) {
        an = A2[i];  // a(L+1) = a(L)
      } else {
        an = alg == 3 ? A1[i] : A0[i];  // a(L+2) = a(L) or a(L-1)
      }
      Am1[i] = A0[i]; A0[i] = A1[i]; A1[i] = A2[i]; A2[i] = an;
      V0[i] = V1[i]; V1[i] = V2[i]; V2[i] = vn;
      ai_prev[i] = ai_j[i];
      dma_j[i] = dma_j1[i];
      pv[i] = qv[i];
      qv[i] = nv[i];
    }
    Xm1 = X0; X0 = X1; X1 = X2; X2 = xn;
    px = qx;
    qx = nx;
  }
  if (k < L) atomicOr(a.error, 2);
}

__global__ void zero_int_kernel(int *p) { *p = 0; }

// a fill kernel rather than a memset node, so the step stays graph-capturable
void launch_zero_int(hipStream_t st, int *p) { hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, st, p); }

void launch_remap(hipStream_t st, const RemapArgs &a) {
  const size_t ncol = (size_t)a.nelem * 16;
  hipLaunchKernelGGL(remap_thickness_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, a);
  // four fields per lane: the column work (grid factors, target search,
  // control flow) is shared four ways at two waves per SIMD
  constexpr int fb = 4;
  const int nfb = (a.nstates + a.qsize + fb - 1) / fb;
  if (nfb == 0) return;
  const int nwaves = ((a.nelem + 3) / 4) * nfb;
  const int nb = (((nwaves + 3) / 4) + 7) & ~7;
  hipLaunchKernelGGL(remap_kernel<fb>, dim3(nb), dim3(256), 0, st, a, nfb);
}

}  // namespace hx
