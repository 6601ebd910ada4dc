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
// that depend on the thicknesses only (the reference's ppmdx rows) and the
// target table are evaluated once per column and shared through LDS by all
// the fields of the element's block.
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

// (the divisions by 3 and 6 below are div_c: the same value, fewer operations)
__device__ __forceinline__ double integrate_parabola(double sq, double lin, double cst, double x1, double x2) {
  return (cst * (x2 - x1) + lin * (x2 * x2 - x1 * x1) / 2.0) + div_c<3>(sq * (x2 * x2 * x2 - x1 * x1 * x1));
}

// dma(j) from a(j-1..j+1) and the grid factors g0, g1, g2 of dx(j-1..j+1)
// (the reference's ppmdx rows 1-3)
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
  double d = ar - al, q = div_c<6>(d * d);
  if (d * (a - (al + ar) / 2.0) > q) {
    al = 3.0 * a - 2.0 * ar;
    d = ar - al;
    q = div_c<6>(d * d);
  }
  if (d * (a - (al + ar) / 2.0) < -q) ar = 3.0 * a - 2.0 * al;
  c0 = 1.5 * a - (al + ar) / 4.0;
  c1 = ar - al;
  c2 = 3.0 * (-2.0 * a + (al + ar));
}

// One block = one element (16 GLL columns) and up to 16*RFB fields; lane
// (row, pt) sweeps column pt for fields f0 + (16*blockIdx.y + row)*RFB + i
// (field 0,1: v components, 2: T, >= 3: tracer field-3).  What depends on
// the column only is computed once per block and shared through LDS:
// * the target table: for each target level k the source cell kid(k) that
//   holds its lower interface and the offset z2(k) in it (the reference's
//   kid and z2 arrays, the same forward search), found by the 16 lanes of
//   row 0 before the sweep;
// * the grid factors of RCH sweep iterations at a time (the reference's
//   ppmdx rows for dma(j+1) and ai(j), dx(j) and the incoming dx(j+3)),
//   iteration j0 + row computed by row row.
constexpr int RCH = 16;  // sweep iterations per grid chunk (rows 0..RCH-1 compute one each)
constexpr int RG = 11;   // shared grid values per (iteration, column)

size_t remap_lds_bytes(int nlev) { return (size_t)RCH * RG * 16 * 8 + (size_t)nlev * (16 * (8 + 8 + 2) + 2 * 8); }

template <int RFB>  // fields per lane
__global__ __launch_bounds__(256) void remap_kernel(RemapArgs a) {
  extern __shared__ double rsm[];
  const int L = a.nlev;
  double *gsh = rsm;                              // [RCH][RG][16]
  double *zsh = rsm + RCH * RG * 16;              // z2  [L][16]
  double *rtk = zsh + (size_t)L * 16;             // 1 / (target thickness)  [L][16]
  double *dab = rtk + (size_t)L * 16;             // hyai(k+1) - hyai(k), hybi(k+1) - hybi(k)  [L][2]
  short *ksh = (short *)(dab + (size_t)L * 2);    // kid [L][16]
  const int ie = xcd_block(blockIdx.x, gridDim.x);
  const int row = threadIdx.x >> 4, pt = threadIdx.x & 15;
  const int nf = 3 + a.qsize;    // in the field numbering above
  const int f0 = 3 - a.nstates;  // rsplit == 0: tracers only
  const int fb = f0 + (blockIdx.y * 16 + row) * RFB;
  const bool active = fb < nf;
  const size_t col = (size_t)ie * 16 + pt;
  const double *src = (a.nstates ? a.dp : a.src) + (size_t)ie * L * 16 + pt;
  const double ps = a.ps[col];
  // wait for that load on every path here (s_waitcnt vmcnt(0); expcnt and
  // lgkmcnt left at their maxima: 0x0F70 on gfx9): only row 0 uses ps before
  // the sweep, and without the wait the compiler's waits inside the sweep
  // assume it may still be in flight and also wait for the prefetched loads
  __builtin_amdgcn_s_waitcnt(0x0F70);
  // dx of Fortran cell m = 1..L with the mirrored ghosts dx(0) = dx(1),
  // dx(-1) = dx(2), dx(L+1) = dx(L), dx(L+2) = dx(L-1)
  auto dxf = [&](int m) {
    int c = m - 1;
    c = c < 0 ? -c - 1 : (c >= L ? 2 * L - 1 - c : c);
    return src[(size_t)c * 16];
  };

  // ------------------------------------------------------------ targets
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    dab[2 * k] = a.hyai[k + 1] - a.hyai[k];
    dab[2 * k + 1] = a.hybi[k + 1] - a.hybi[k];
  }
  // The element's source thicknesses go through LDS (the grid chunk's space,
  // free until the sweep) when they fit, so the serial search below waits on
  // LDS rather than on one global load per step.
  const bool lsrc = L <= RCH * RG;
  if (lsrc) {
    const double *s0 = (a.nstates ? a.dp : a.src) + (size_t)ie * L * 16;
    for (int i = threadIdx.x; i < L * 16; i += blockDim.x) gsh[i] = s0[i];
  }
  __syncthreads();
  auto search = [&](auto sat) {  // sat(c): thickness of source cell c of this column
    double pio_L = 0.0;  // pio(L): serial sum of the source thicknesses
    for (int c = 0; c < L; ++c) pio_L += sat(c);
    double pin_acc = 0.0, pio_si = 0.0, pio_m1 = 0.0, pio_m2 = 0.0;  // pio(si), pio(si-1), pio(si-2)
    int si = 0;
    auto advance_si = [&]() {
      pio_m2 = pio_m1;
      pio_m1 = pio_si;
      pio_si = si < L ? pio_si + sat(si) : pio_si + 1.0;  // pio(L+1) = pio(L) + 1
      ++si;
    };
    for (int k = 0; k < L; ++k) {
      double pin_k1;
      if (k + 1 < L) {
        pin_acc += dab[2 * k] * a.ps0 + dab[2 * k + 1] * ps;  // the target thickness of level k
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
      const int kid = kk - 1;
      ksh[k * 16 + pt] = (short)kid;
      zsh[k * 16 + pt] = kid >= 0 ? (pin_k1 - (plo + phi) * 0.5) / sat(kid) : 0.0;
    }
  };
  if (row == 0) {
    if (lsrc) search([&](int c) { return gsh[c * 16 + pt]; });
    else search([&](int c) { return src[(size_t)c * 16]; });
  } else {
    // meanwhile the other rows form the reciprocals of the target
    // thicknesses (the outputs' divisor, shared by the fields)
    for (int k = row - 1; k < L; k += 15) rtk[k * 16 + pt] = 1.0 / (dab[2 * k] * a.ps0 + dab[2 * k + 1] * ps);
  }

  // ------------------------------------------------------------ fields
  // (pointers of absent fields stay valid device addresses, never null, so
  // every access compiles to a global, not a flat, instruction).  An absent
  // field slot of a lane (on[i] false: past the last tracer) is swept like
  // the others on its stand-in values and only its stores are skipped: the
  // sweep then has no per-field branches (1941 -> 1891 us at 43 fields)
  // (its loads read the first L*16 values of T, in bounds)
  double *var[RFB];
  double *Qout[RFB];
  size_t vs[RFB];
  bool on[RFB], state[RFB], qout[RFB];
#pragma unroll
  for (int i = 0; i < RFB; ++i) {
    const int f = fb + i;
    on[i] = f < nf;
    state[i] = f < 3;
    qout[i] = f >= 3 && f < nf && a.Q != nullptr;
    vs[i] = f < 2 ? 32 : 16;
    var[i] = a.t;
    Qout[i] = a.t;
    if (f < 2) var[i] = a.v + (size_t)ie * L * 32 + f * 16 + pt;
    else if (f == 2) var[i] = a.t + (size_t)ie * L * 16 + pt;
    else if (f < nf) {
      var[i] = a.qdp + (((size_t)ie * a.qsize + (f - 3)) * L) * 16 + pt;
      if (qout[i]) Qout[i] = a.Q + (((size_t)ie * a.qsize + (f - 3)) * L) * 16 + pt;
    }
  }
  const int alg = a.remap_alg;
  // Fortran-style cell indices j = 1..L.  Per field (A0..A2) = a(j..j+2),
  // V0..V2 = var(j..j+2)
  double A0[RFB], A1[RFB], A2[RFB], V0[RFB], V1[RFB], V2[RFB];
  double dma_j[RFB], ai_prev[RFB], var_prefix[RFB], mass_prev[RFB];
  double cc0[RFB], cc1[RFB], cc2[RFB];
  double pv[RFB], qv[RFB];  // field values of cells j+3, j+4 (loaded ahead)
  // unconditional loads (past the bottom the last cell is reloaded; an
  // absent field reads its stand-in pointer): with a branch here every later
  // wait would also cover the loads just issued
  auto fetch = [&](int m, double *v) {
    const size_t c = (size_t)((m <= L ? m : L) - 1);
#pragma unroll
    for (int i = 0; i < RFB; ++i) v[i] = var[i][c * vs[i]];
  };
  if (active) {
    const double s0 = src[0], s1 = src[16];
    // dma(0) factors: dx(-1), dx(0), dx(1) = s1, s0, s0
    const double g0 = s0 / (s1 + s0 + s0);
    const double g1 = (2.0 * s1 + s0) / (s0 + s0);
    const double g2 = (s0 + 2.0 * s0) / (s1 + s0);
#pragma unroll
    for (int i = 0; i < RFB; ++i) {
      const double x0 = var[i][0], x1 = var[i][vs[i]];
      const double v0 = state[i] ? x0 * s0 : x0, v1 = state[i] ? x1 * s1 : x1;
      const double a1 = v0 / s0, a2 = v1 / s1;
      const double am1 = alg == 3 ? a1 : a2;  // a(-1)
      A0[i] = a1;  // a(0) = a(1)
      A1[i] = a1;
      A2[i] = a2;
      V0[i] = 0.0;
      V1[i] = v0;
      V2[i] = v1;
      dma_j[i] = ppm_dma_g(am1, A0[i], A1[i], g0, g1, g2);  // dma(0)
      ai_prev[i] = 0.0;
      var_prefix[i] = 0.0;
      mass_prev[i] = 0.0;
    }
    fetch(3, pv);
    fetch(4, qv);
  }

  // ------------------------------------------------------------ sweep
  int k = 0;  // next target level
  bool live = active;
  for (int j0 = 0; j0 <= L; j0 += RCH) {
    __syncthreads();  // the previous chunk's grid values are consumed (and the target table is written)
    {
      const int j = j0 + row;
      if (row < RCH && j <= L) {
        // window dx(j-1..j+2): dma(j+1) factors use cells j..j+2, ai(j) factors cells j-1..j+2
        const double Xm1 = dxf(j - 1), X0 = dxf(j), X1 = dxf(j + 1), X2 = dxf(j + 2);
        double *g = gsh + row * RG * 16 + pt;
        g[0] = X1 / (X0 + X1 + X2);
        g[16] = (2.0 * X0 + X1) / (X1 + X2);
        g[32] = (X1 + 2.0 * X2) / (X0 + X1);
        g[48] = X0 / (X0 + X1);
        g[64] = 1.0 / (Xm1 + X0 + X1 + X2);
        g[80] = (2.0 * X0 * X1) / (X0 + X1) * ((Xm1 + X0) / (2.0 * X0 + X1) - (X2 + X1) / (2.0 * X1 + X0));
        g[96] = X0 * (Xm1 + X0) / (2.0 * X0 + X1);
        g[112] = X1 * (X1 + X2) / (X0 + 2.0 * X1);
        g[128] = X0;
        g[144] = dxf(j + 3);  // the cell entering the window (used for j + 3 <= L)
        g[160] = 1.0 / g[144];
      }
    }
    __syncthreads();
    if (!live) continue;
    const int jn = L + 1 - j0 < RCH ? L + 1 - j0 : RCH;
    // (unrolled by 3: the window's shift through A0..A2 and V0..V2 becomes
    // register renaming in two of three steps; 2030 -> 2004 us)
#pragma unroll 3
    for (int jj = 0; jj < jn; ++jj) {
      const int j = j0 + jj;
      double nv[RFB];
      fetch(j + 5, nv);
      const double *g = gsh + jj * RG * 16 + pt;
      const double h0 = g[0], h1 = g[16], h2 = g[32];
      const double gai[5] = {g[48], g[64], g[80], g[96], g[112]};
      const double X0 = g[128];
      double dma_j1[RFB], ai_j[RFB];
#pragma unroll
      for (int i = 0; i < RFB; ++i) {
        dma_j1[i] = ppm_dma_g(A0[i], A1[i], A2[i], h0, h1, h2);       // dma(j+1)
        ai_j[i] = ppm_ai_g(A0[i], A1[i], dma_j[i], dma_j1[i], gai);  // ai(j)
      }
      if (j >= 1) {
        const int c = j - 1;  // physical source cell, Fortran cell j
#pragma unroll
        for (int i = 0; i < RFB; ++i) {
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
          const int kid = ksh[k * 16 + pt];
          if (kid > c) break;
          if (kid < c) {
            atomicOr(a.error, 2);  // deformation beyond one layer: not supported
            live = false;
            break;
          }
          const double z2 = zsh[k * 16 + pt];
          const double tk = dab[2 * k] * a.ps0 + dab[2 * k + 1] * ps;  // the reference's (hyai(k+1)-hyai(k))*ps0 + ...
          const Recip rt{tk, rtk[k * 16 + pt]};
#pragma unroll
          for (int i = 0; i < RFB; ++i) {
            const double integral = integrate_parabola(cc2[i], cc1[i], cc0[i], -0.5, z2);
            const double mass = var_prefix[i] + integral * X0;
            const double out = k > 0 ? mass - mass_prev[i] : mass;
            mass_prev[i] = mass;
            if (!on[i]) continue;
            if (state[i]) {
              var[i][(size_t)k * vs[i]] = div_r(out, rt);
            } else {
              var[i][(size_t)k * vs[i]] = out;
              if (qout[i]) ((__attribute__((address_space(1))) double *)Qout[i])[(size_t)k * 16] = div_r(out, rt);
            }
          }
          ++k;
        }
        if (!live) break;
#pragma unroll
        for (int i = 0; i < RFB; ++i) var_prefix[i] += V0[i];
      }
      // shift: bring in cell j+3
      const int m = j + 3;
      const Recip rxn{g[144], g[160]};
#pragma unroll
      for (int i = 0; i < RFB; ++i) {
        double an, vn = 0.0;
        if (m <= L) {
          const double x = pv[i];
          vn = state[i] ? x * rxn.b : x;
          an = div_r(vn, rxn);
        } else if (m == L + 1) {
          an = A2[i];  // a(L+1) = a(L)
        } else {
          an = alg == 3 ? A1[i] : A0[i];  // a(L+2) = a(L) or a(L-1)
        }
        A0[i] = A1[i]; A1[i] = A2[i]; A2[i] = an;
        V0[i] = V1[i]; V1[i] = V2[i]; V2[i] = vn;
        ai_prev[i] = ai_j[i];
        dma_j[i] = dma_j1[i];
        pv[i] = qv[i];
        qv[i] = nv[i];
      }
    }
  }
  if (live && k < L) atomicOr(a.error, 2);
}

__global__ void zero_int_kernel(int *p) { *p = 0; }

// a fill kernel rather than a memset node, so the step stays graph-capturable
void launch_zero_int(hipStream_t st, int *p) { hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, st, p); }

// ps_v = sum(dp3d) + hyai0*ps0, summed as remap_thickness_kernel does
__global__ __launch_bounds__(256) void ps_from_dp_kernel(int nelem, int nlev, double hyai0, double ps0,
                                                          const double *dp, double *ps) {
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (size_t)nelem * 16) return;
  const double *d = dp + (col >> 4) * nlev * 16 + (col & 15);
  double s = 0.0;
  for (int k = 0; k < nlev; ++k) s += d[(size_t)k * 16];
  ps[col] = s + hyai0 * ps0;
}

void launch_ps_from_dp(hipStream_t st, int nelem, int nlev, double hyai0, double ps0, const double *dp, double *ps) {
  const size_t ncol = (size_t)nelem * 16;
  hipLaunchKernelGGL(ps_from_dp_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, nelem, nlev, hyai0, ps0, dp, ps);
}

void launch_remap(hipStream_t st, const RemapArgs &a) {
  const size_t ncol = (size_t)a.nelem * 16;
  hipLaunchKernelGGL(remap_thickness_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, a);
  const int nfld = a.nstates + a.qsize;
  if (nfld == 0) return;
  // fields per lane: as few as fill the 16 rows (at most 4, then more
  // blocks per element)
  const int rfb = nfld <= 16 ? 1 : nfld <= 32 ? 2 : nfld <= 48 ? 3 : 4;
  const dim3 g(a.nelem, (nfld + 16 * rfb - 1) / (16 * rfb));
  const size_t lds = remap_lds_bytes(a.nlev);
  switch (rfb) {
    case 1: hipLaunchKernelGGL(remap_kernel<1>, g, dim3(256), lds, st, a); break;
    case 2: hipLaunchKernelGGL(remap_kernel<2>, g, dim3(256), lds, st, a); break;
    case 3: hipLaunchKernelGGL(remap_kernel<3>, g, dim3(256), lds, st, a); break;
    default: hipLaunchKernelGGL(remap_kernel<4>, g, dim3(256), lds, st, a); break;
  }
}

}  // namespace hx
