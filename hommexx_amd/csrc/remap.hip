// Vertical remap of (v dp, T dp, Qdp) from the Lagrangian levels back to the
// reference hybrid levels with monotone PPM (RemapFunctor.hpp:226-331,
// PpmRemap.hpp:78-673 in the reference), plus update_q (prim_driver.cpp:171-206).
//
// One lane = one (GLL column, field).  The reference builds per-column arrays
// (pio, pin, ppmdx[10], kid, z2, ao, dma, ai, parabola coefficients) and then
// remaps; here one top-to-bottom sweep produces the same numbers with O(1)
// state: source cell j's parabola needs ai(j-1), ai(j), which need
// dma(j-1..j+1) and a, dx up to j+2, so a four-cell window suffices, and the
// target interface k+1 lies in source cell kid[k] >= k-1 (kid is found by the
// same forward search).  Every value is computed with the reference's
// expression and operand order.
#include "kernels.hpp"

namespace hx {

// ps_v = sum(dp3d) + hyai0*ps0 and the thickness check (ComputeThicknessTag)
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
    bad |= isnan(x) || x < 0.0;
  }
  s += a.hyai0 * a.ps0;
  a.ps[col] = s;
  if (bad) atomicOr(a.error, 1);
}

struct Cell {
  double c0, c1, c2;  // parabola: const, linear, square
  double mass;        // exclusive prefix of the remapped quantity at this cell
  double dx;          // source thickness
};

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

// ai(j) from a(j), a(j+1), dma(j), dma(j+1), dx(j-1..j+2)
__device__ __forceinline__ double ppm_ai(double a0, double a1, double dma0, double dma1, double xm1, double x0,
                                         double x1, double x2) {
  const double g3 = x0 / (x0 + x1);
  const double g4 = 1.0 / (xm1 + x0 + x1 + x2);
  const double g5 = (2.0 * x0 * x1) / (x0 + x1);
  const double g6 = (xm1 + x0) / (2.0 * x0 + x1);
  const double g7 = (x2 + x1) / (2.0 * x1 + x0);
  const double g8 = x0 * (xm1 + x0) / (2.0 * x0 + x1);
  const double g9 = x1 * (x1 + x2) / (x0 + 2.0 * x1);
  return a0 + g3 * (a1 - a0) + g4 * (g5 * (g6 - g7) * (a1 - a0) - g8 * dma1 + g9 * dma0);
}

__device__ __forceinline__ void ppm_parabola(double al, double ar, double a, double &c0, double &c1, double &c2) {
  if ((ar - a) * (a - al) <= 0.) {
    al = a;
    ar = a;
  }
  if ((ar - al) * (a - (al + ar) / 2.0) > (ar - al) * (ar - al) / 6.0) al = 3.0 * a - 2.0 * ar;
  if ((ar - al) * (a - (al + ar) / 2.0) < -(ar - al) * (ar - al) / 6.0) ar = 3.0 * a - 2.0 * al;
  c0 = 1.5 * a - (al + ar) / 4.0;
  c1 = ar - al;
  c2 = 3.0 * (-2.0 * a + (al + ar));
}

// One (column, field).  field 0,1: v components, 2: T, >= 3: tracer field-3.
__global__ __launch_bounds__(256) void remap_kernel(RemapArgs a) {
  const int nf = 3 + a.qsize;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // 16 lanes = the 16 columns of one element for one field, so each level is
  // one coalesced 128-byte access; the fields of an element follow each other
  // (the element's dp3d columns stay in cache across its fields)
  const int pt = (int)(tid & 15);
  const size_t rest = tid >> 4;
  const int f = (int)(rest % nf);
  const size_t ie_s = rest / nf;
  if (ie_s >= (size_t)a.nelem) return;
  const int ie = (int)ie_s;
  const size_t col = (size_t)ie * 16 + pt;
  const int L = a.nlev;
  const double *src = a.dp + (size_t)ie * L * 16 + pt;
  double *var;
  size_t vs = 16;
  if (f < 2) {
    var = a.v + (size_t)ie * L * 32 + f * 16 + pt;
    vs = 32;
  } else if (f == 2) {
    var = a.t + (size_t)ie * L * 16 + pt;
  } else {
    var = a.qdp + (((size_t)ie * a.qsize + (f - 3)) * L) * 16 + pt;
  }
  double *Qout = (f >= 3 && a.Q) ? a.Q + (((size_t)ie * a.qsize + (f - 3)) * L) * 16 + pt : nullptr;
  const bool state = f < 3;
  const double ps = a.ps[col];
  const int alg = a.remap_alg;
  auto tgt_of = [&](int k) {
    return (a.hyai[k + 1] - a.hyai[k]) * a.ps0 + (a.hybi[k + 1] - a.hybi[k]) * ps;
  };
  // extensive value of physical cell c (states are multiplied by dp3d first)
  auto load_var = [&](int c, double dxc) {
    const double x = var[(size_t)c * vs];
    return state ? x * dxc : x;
  };
  // pio(L): serial sum of the source thicknesses
  double pio_L = 0.0;
  for (int c = 0; c < L; ++c) pio_L += src[(size_t)c * 16];

  // ------------------------------------------------------------ window
  // Fortran-style cell indices j = 1..L, mirrored (or constant) ghosts.
  // (Xm1, X0, X1, X2) = dx(j-1..j+2), (Am1..A2) = a(j-1..j+2), V0..V2 = var(j..j+2)
  const double s0 = src[0], s1 = src[16];
  const double v0 = load_var(0, s0), v1 = load_var(1, s1);
  const double a1 = v0 / s0, a2 = v1 / s1;
  double Xm1 = s1, X0 = s0, X1 = s0, X2 = s1;
  double Am1 = alg == 3 ? a1 : a2, A0 = a1, A1 = a1, A2 = a2;
  double V0 = 0.0, V1 = v0, V2 = v1;  // var(0) unused
  double dma_j = ppm_dma(Am1, A0, A1, Xm1, X0, X1);  // dma(0)
  double ai_prev = 0.0;
  double var_prefix = 0.0;  // sum of var over physical cells < c

  // ------------------------------------------------------------ targets
  int k = 0;                 // next target level
  int kid = -2;              // kid of target k (-2: not yet computed)
  double z2 = 0.0;
  double pin_acc = 0.0;      // serial prefix of the target thicknesses
  int si = 0;                // next source interface to test
  double pio_si = 0.0, pio_m1 = 0.0, pio_m2 = 0.0;  // pio(si), pio(si-1), pio(si-2)
  double mass_prev = 0.0;
  Cell prev{0, 0, 0, 0, 0};
  bool have_prev = false;

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
  auto finish = [&](const Cell &cc) {
    const double integral = integrate_parabola(cc.c2, cc.c1, cc.c0, -0.5, z2);
    const double mass = cc.mass + integral * cc.dx;
    const double out = k > 0 ? mass - mass_prev : mass;
    mass_prev = mass;
    const double tk = tgt_of(k);
    if (state) {
      var[(size_t)k * vs] = out / tk;
    } else {
      var[(size_t)k * vs] = out;
      if (Qout) Qout[(size_t)k * 16] = out / tk;
    }
    ++k;
    kid = -2;
  };

  // ------------------------------------------------------------ sweep
  for (int j = 0; j <= L; ++j) {
    const double dma_j1 = ppm_dma(A0, A1, A2, X0, X1, X2);               // dma(j+1)
    const double ai_j = ppm_ai(A0, A1, dma_j, dma_j1, Xm1, X0, X1, X2);  // ai(j)
    if (j >= 1) {
      const int c = j - 1;  // physical source cell, Fortran cell j
      Cell cc;
      ppm_parabola(ai_prev, ai_j, A0, cc.c0, cc.c1, cc.c2);
      if (alg == 2 && (c <= 1 || c >= L - 2)) {
        cc.c0 = A0;
        cc.c1 = 0.0;
        cc.c2 = 0.0;
      }
      cc.dx = X0;
      cc.mass = var_prefix;
      while (k < L) {
        if (kid == -2) compute_kid();
        if (kid == c) {
          finish(cc);
        } else if (kid == c - 1 && have_prev) {
          finish(prev);
        } else if (kid < c) {
          atomicOr(a.error, 2);  // deformation beyond one layer: not supported
          return;
        } else {
          break;
        }
      }
      prev = cc;
      have_prev = true;
      var_prefix += V0;
    }
    // shift: bring in cell j+3
    const int m = j + 3;
    double xn, an, vn = 0.0;
    if (m <= L) {
      xn = src[(size_t)(m - 1) * 16];
      vn = load_var(m - 1, xn);
      an = vn / xn;
    } else if (m == L + 1) {
      xn = X2;  // dx(L+1) = dx(L)
      an = A2;  // a(L+1) = a(L)
    } else {
      xn = X0;                 // dx(L+2) = dx(L-1)
      an = alg == 3 ? A1 : A0;  // a(L+2) = a(L) or a(L-1)
    }
    Xm1 = X0; X0 = X1; X1 = X2; X2 = xn;
    Am1 = A0; A0 = A1; A1 = A2; A2 = an;
    V0 = V1; V1 = V2; V2 = vn;
    ai_prev = ai_j;
    dma_j = dma_j1;
  }
  if (k < L) atomicOr(a.error, 2);
}

__global__ void zero_int_kernel(int *p) { *p = 0; }

// a fill kernel rather than a memset node, so the step stays graph-capturable
void launch_zero_int(hipStream_t st, int *p) { hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, st, p); }

void launch_remap(hipStream_t st, const RemapArgs &a) {
  const size_t ncol = (size_t)a.nelem * 16;
  hipLaunchKernelGGL(remap_thickness_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, a);
  const size_t n = ncol * (3 + a.qsize);
  hipLaunchKernelGGL(remap_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace hx
