// Biharmonic hyperviscosity on the np1 state, one subcycle per call
// (HyperviscosityFunctorImpl.cpp:56-112, HyperviscosityFunctorImpl.hpp:71-257).
// Slab kernels: a 16-lane group = one element at one level.
//   lap1      : weak laplacians of T, dp3d and the vector laplacian of v
//   DSS(*rsmp)
//   lap2+pre  : second laplacians, dpdiss accumulators, -nu scaling, the
//               nu_top sponge on the top three levels, dptens <- dp3d*sph + dt*dptens
//   DSS + update : assemble, apply the inverse mass, v += , T += - heating/cp,
//                  dp3d = dptens*rsmp
// The mixed form swaps in FP32 laplacian kernels (below) and an FP32 first
// exchange; the second exchange and the update are the same.
#include "kernels.hpp"

namespace hx {

__device__ __forceinline__ bool slab_of(int nblk, int nelem, int nlev, int &ie, int &k, int &pt) {
  const int b = xcd_block(blockIdx.x, nblk);
  const size_t slab = (size_t)b * 16 + (threadIdx.x >> 4);
  pt = threadIdx.x & 15;
  if (slab >= (size_t)nelem * nlev) return false;
  ie = (int)(slab / nlev);
  k = (int)(slab % nlev);
  return true;
}

__global__ __launch_bounds__(256) void hv_lap1_kernel(HvArgs a) {
  int ie, k, pt;
  if (!slab_of(gridDim.x, a.nelem, a.nlev, ie, k, pt)) return;
  const size_t s = ((size_t)ie * a.nlev + k) * 16 + pt;
  const size_t v = ((size_t)ie * a.nlev + k) * 32 + pt;
  hv_first_laplacians<false>(a.m, ie, pt, a.nu_ratio1, a.t[s], a.dp[s], a.v[v], a.v[v + 16], a.tt_a, a.dpt_a, a.vt_a,
                             s, v);
}

// The first exchange is fused in: each lane gathers the assembled first
// laplacian of its point (own value and neighbour copies of the vt_a, tt_a,
// dpt_a outputs of hv_lap1, in dss_kernel's order, times rspheremp) and the
// outputs go to vt_b, tt_b, dpt_b, which hv_dss_update then exchanges (the
// inputs of a neighbour's gather are never overwritten during the launch).
// Slabs are walked in level groups (group_slab) so neighbour copies come
// from L2.
// TOP: the nu_top sponge levels 0..2 only (a launch of their own, so that the
// sponge's extra laplacians do not raise the register count of the main
// launch: 123 VGPRs without them, 152 with); !TOP with nu_top > 0 skips them.
template <bool TENSOR, bool TOP>
__global__ __launch_bounds__(256) void hv_lap2_pre_kernel(HvArgs a) {
  int ie, k, pt;
  if (!group_slab(a.nelem, TOP ? (a.nlev < 3 ? a.nlev : 3) : a.nlev, ie, k, pt)) return;
  if (!TOP && a.nu_top > 0 && k < 3) return;  // whole slab
  const size_t s = ((size_t)ie * a.nlev + k) * 16 + pt;
  const size_t v = ((size_t)ie * a.nlev + k) * 32 + pt;
  double g[4];  // assembled vt (2 components), tt, dpt
  {
    const int *src = a.src + ((size_t)ie * 16 + pt) * 3;
    const int s0 = src[0], s1 = src[1], s2 = src[2];
    const size_t es = (size_t)a.nlev * 16, ev = (size_t)a.nlev * 32;
    auto ld = [&](const double *f, size_t estr, size_t lstr, size_t c, double (&r)[4]) {
      const size_t off = (size_t)k * lstr + c, own = (size_t)ie * estr + off + pt;
      auto at = [&](int sr) { return opaque_index(sr >= 0 ? (size_t)(sr >> 4) * estr + off + (sr & 15) : own); };
      r[0] = f[own];
      r[1] = f[at(s0)];
      r[2] = f[at(s1)];
      r[3] = f[at(s2)];
    };
    double r0[4], r1[4], rt[4], rdp[4];
    ld(a.vt_a, ev, 32, 0, r0);
    ld(a.vt_a, ev, 32, 16, r1);
    ld(a.tt_a, es, 16, 0, rt);
    ld(a.dpt_a, es, 16, 0, rdp);
    const double rs = a.m.rspheremp[(size_t)ie * 16 + pt];
    g[0] = gather_sum(r0, s0, s1, s2) * rs;
    g[1] = gather_sum(r1, s0, s1, s2) * rs;
    g[2] = gather_sum(rt, s0, s1, s2) * rs;
    g[3] = gather_sum(rdp, s0, s1, s2) * rs;
  }
  ElemLane e;
  load_elem(a.m, ie, pt, e);
  DvvLane d;
  load_dvv(a.m.dvv, pt, d);
  // second laplacian: constant coefficient, or the tensor form
  // (HyperviscosityFunctorImpl.hpp:88-128), then the pre-exchange scaling
  // with the nu_top sponge on the top three levels; each field is finished
  // and stored before the next one starts (fewer live registers)
  double tv[4], vs[6];
  if (TENSOR) load_tensor(a.m, ie, pt, tv, vs);
  const bool top = TOP;  // (the TOP launch runs only with nu_top > 0)
  const double sc = (k == 0 ? 4.0 : k == 1 ? 2.0 : 1.0) * a.nu_top;
  {
    double vt0, vt1;
    if (TENSOR) sph_vlaplace_cartesian(e, d, tv, vs, g[0], g[1], vt0, vt1);
    else sph_vlaplace_contra(e, d, a.nu_ratio2, g[0], g[1], vt0, vt1);
    vt0 *= -a.nu;
    vt1 *= -a.nu;
    if (top) {
      double lv0, lv1;
      sph_vlaplace_contra(e, d, 1.0, a.v[v], a.v[v + 16], lv0, lv1);
      vt0 += sc * lv0;
      vt1 += sc * lv1;
    }
    a.vt_b[v] = vt0;
    a.vt_b[v + 16] = vt1;
  }
  {
    double tt = TENSOR ? sph_laplace_tensor(e, d, tv, g[2]) : sph_laplace(e, d, g[2]);
    tt *= -a.nu_s;
    if (top) tt += sc * sph_laplace(e, d, a.t[s]);
    a.tt_b[s] = tt;
  }
  const double dpn = a.dp[s];
  double dpt = TENSOR ? sph_laplace_tensor(e, d, tv, g[3]) : sph_laplace(e, d, g[3]);
  a.dpdiss_ave[s] += a.eta_ave_w * dpn / a.nsub;
  a.dpdiss_bih[s] += a.eta_ave_w * dpt / a.nsub;
  dpt *= -a.nu_p;
  if (top) dpt += sc * sph_laplace(e, d, dpn);
  dpt *= a.dt;
  dpt += dpn * e.sph;
  a.dpt_b[s] = dpt;
}

// ---------------------------------------------------------- mixed form
// The two laplacians in FP32 (compensated contractions).  The first
// laplacian's outputs and their assembled values are FP32 arrays held in the
// first halves of vt_a/tt_a/dpt_a and vt_b/tt_b/dpt_b; T and dp3d are shifted
// by a per-slab reference value before rounding (laplacians annihilate
// constants).  Accumulators, scaling, the sponge combination and everything
// after stay FP64.
__global__ __launch_bounds__(256) void hv_lap1_mixed_kernel(HvArgs a) {
  int ie, k, pt;
  if (!slab_of(gridDim.x, a.nelem, a.nlev, ie, k, pt)) return;
  const size_t s = ((size_t)ie * a.nlev + k) * 16 + pt;
  const size_t v = ((size_t)ie * a.nlev + k) * 32 + pt;
  hv_first_laplacians<true>(a.m, ie, pt, a.nu_ratio1, a.t[s], a.dp[s], a.v[v], a.v[v + 16], a.tt_a, a.dpt_a, a.vt_a,
                            s, v);
}

// As hv_lap2_pre: the first exchange of the FP32 first laplacian is gathered
// here (summed in FP64, times rspheremp, rounded to FP32 like the exchanged
// array) and the outputs go to the _b buffers.
__global__ __launch_bounds__(256) void hv_lap2_pre_mixed_kernel(HvArgs a) {
  int ie, k, pt;
  if (!group_slab(a.nelem, a.nlev, ie, k, pt)) return;
  const size_t s = ((size_t)ie * a.nlev + k) * 16 + pt;
  const size_t v = ((size_t)ie * a.nlev + k) * 32 + pt;
  float g[4];  // assembled vt (2 components), tt, dpt
  {
    const int *src = a.src + ((size_t)ie * 16 + pt) * 3;
    const int s0 = src[0], s1 = src[1], s2 = src[2];
    const size_t es = (size_t)a.nlev * 16, ev = (size_t)a.nlev * 32;
    auto ld = [&](const double *fd, size_t estr, size_t lstr, size_t c, float (&r)[4]) {
      const float *f = (const float *)fd;
      const size_t off = (size_t)k * lstr + c, own = (size_t)ie * estr + off + pt;
      auto at = [&](int sr) { return opaque_index(sr >= 0 ? (size_t)(sr >> 4) * estr + off + (sr & 15) : own); };
      r[0] = f[own];
      r[1] = f[at(s0)];
      r[2] = f[at(s1)];
      r[3] = f[at(s2)];
    };
    float r0[4], r1[4], rt[4], rdp[4];
    ld(a.vt_a, ev, 32, 0, r0);
    ld(a.vt_a, ev, 32, 16, r1);
    ld(a.tt_a, es, 16, 0, rt);
    ld(a.dpt_a, es, 16, 0, rdp);
    const double rs = a.m.rspheremp[(size_t)ie * 16 + pt];
    g[0] = (float)(gather_sum(r0, s0, s1, s2) * rs);
    g[1] = (float)(gather_sum(r1, s0, s1, s2) * rs);
    g[2] = (float)(gather_sum(rt, s0, s1, s2) * rs);
    g[3] = (float)(gather_sum(rdp, s0, s1, s2) * rs);
  }
  ElemLaneF e;
  load_elemf(a.m, ie, pt, e);
  DvvLaneF d;
  load_dvvf(a.m.dvv, pt, d);
  const double sph = a.m.spheremp[(size_t)ie * 16 + pt];
  double tt = sph_laplacef(e, d, g[2]);
  double dpt = sph_laplacef(e, d, g[3]);
  float f0, f1;
  sph_vlaplace_contraf(e, d, (float)a.nu_ratio2, g[0], g[1], f0, f1);
  double vt0 = f0, vt1 = f1;
  const double dpn = a.dp[s];
  a.dpdiss_ave[s] += a.eta_ave_w * dpn / a.nsub;
  a.dpdiss_bih[s] += a.eta_ave_w * dpt / a.nsub;
  double lv0 = 0, lv1 = 0, lt = 0, ldp = 0;
  const bool top = a.nu_top > 0 && k < 3;
  if (top) {
    float g0, g1;
    sph_vlaplace_contraf(e, d, 1.f, (float)a.v[v], (float)a.v[v + 16], g0, g1);
    lv0 = g0;
    lv1 = g1;
    const double t = a.t[s];
    lt = sph_laplacef(e, d, (float)(t - slab_ref(t)));
    ldp = sph_laplacef(e, d, (float)(dpn - slab_ref(dpn)));
  }
  vt0 *= -a.nu;
  vt1 *= -a.nu;
  tt *= -a.nu_s;
  dpt *= -a.nu_p;
  if (top) {
    const double sc = (k == 0 ? 4.0 : k == 1 ? 2.0 : 1.0) * a.nu_top;
    vt0 += sc * lv0;
    vt1 += sc * lv1;
    tt += sc * lt;
    dpt += sc * ldp;
  }
  dpt *= a.dt;
  dpt += dpn * sph;
  a.vt_b[v] = vt0;
  a.vt_b[v + 16] = vt1;
  a.tt_b[s] = tt;
  a.dpt_b[s] = dpt;
}

// Second exchange fused with the update: the assembled tensors are used at
// once by the lane that assembled them (same gather order as dss_kernel).
// LAP1 (1: FP64, 2: mixed form): the same kernel also forms the first
// laplacians of the updated state for the next subcycle (what hv_lap1_kernel
// or hv_lap1_mixed_kernel would compute from the stored values), from the
// values in registers; they go to vt_b, tt_b, dpt_b, which the caller points
// at the next subcycle's first-laplacian buffers.
template <int LAP1>
__global__ __launch_bounds__(256) void hv_dss_update_kernel(HvArgs a) {
  int ie, k, pt;
  if (!group_slab(a.nelem, a.nlev, ie, k, pt)) return;
  const int *src = a.src + ((size_t)ie * 16 + pt) * 3;
  const int s0 = src[0], s1 = src[1], s2 = src[2];
  const size_t es = (size_t)a.nlev * 16, ev = (size_t)a.nlev * 32;
  // the loads of the four gathers are issued before the first sum
  auto ld = [&](const double *f, size_t estr, size_t lstr, size_t c, double (&r)[4]) {
    const size_t off = (size_t)k * lstr + c, own = (size_t)ie * estr + off + pt;
    auto at = [&](int src) { return opaque_index(src >= 0 ? (size_t)(src >> 4) * estr + off + (src & 15) : own); };
    r[0] = f[own];
    r[1] = f[at(s0)];
    r[2] = f[at(s1)];
    r[3] = f[at(s2)];
  };
  double r0[4], r1[4], rt[4], rdp[4];
  ld(a.vt_a, ev, 32, 0, r0);
  ld(a.vt_a, ev, 32, 16, r1);
  ld(a.tt_a, es, 16, 0, rt);
  ld(a.dpt_a, es, 16, 0, rdp);
  const double b0 = gather_sum(r0, s0, s1, s2), b1 = gather_sum(r1, s0, s1, s2);
  const double bt = gather_sum(rt, s0, s1, s2), bdp = gather_sum(rdp, s0, s1, s2);
  const double r = a.m.rspheremp[(size_t)ie * 16 + pt];
  const size_t i = ((size_t)ie * a.nlev + k) * 16 + pt;
  const size_t v0 = ((size_t)ie * a.nlev + k) * 32 + pt, v1 = v0 + 16;
  const double vt0 = (a.dt * b0 * r);
  const double vt1 = (a.dt * b1 * r);
  const double u = a.v[v0] + vt0, w = a.v[v1] + vt1;
  a.v[v0] = u;
  a.v[v1] = w;
  const double tt = (a.dt * bt * r);
  const double heating = vt0 * u + vt1 * w;
  const double tn = a.t[i] + tt - heating / CP, dpn = (bdp * r);
  a.t[i] = tn;
  a.dp[i] = dpn;
  if (LAP1) hv_first_laplacians<LAP1 == 2>(a.m, ie, pt, a.nu_ratio1, tn, dpn, u, w, a.tt_b, a.dpt_b, a.vt_b, i, v0);
}

void launch_hv_cycle(hipStream_t st, const HvArgs &a, bool lap1_done, bool lap1_next) {
  const size_t nslab = (size_t)a.nelem * a.nlev;
  const int blocks = (int)((nslab + 15) / 16);
  const dim3 gg(group_blocks_x(a.nelem, a.nlev));
  // the first exchange is gathered by hv_lap2_pre, which writes the second
  // laplacian to the _b buffers; the update exchanges those (and, FP64 form,
  // may form the next subcycle's first laplacians on the way)
  if (!lap1_done) {
    if (a.mixed) hipLaunchKernelGGL(hv_lap1_mixed_kernel, dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(hv_lap1_kernel, dim3(blocks), dim3(256), 0, st, a);
  }
  if (a.mixed) hipLaunchKernelGGL(hv_lap2_pre_mixed_kernel, gg, dim3(256), 0, st, a);
  else if (a.consthv) hipLaunchKernelGGL((hv_lap2_pre_kernel<false, false>), gg, dim3(256), 0, st, a);
  else hipLaunchKernelGGL((hv_lap2_pre_kernel<true, false>), gg, dim3(256), 0, st, a);
  if (!a.mixed && a.nu_top > 0) {
    const dim3 gt(group_blocks_x(a.nelem, a.nlev < 3 ? a.nlev : 3));
    if (a.consthv) hipLaunchKernelGGL((hv_lap2_pre_kernel<false, true>), gt, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((hv_lap2_pre_kernel<true, true>), gt, dim3(256), 0, st, a);
  }
  HvArgs b = a;
  b.vt_a = a.vt_b;
  b.tt_a = a.tt_b;
  b.dpt_a = a.dpt_b;
  b.vt_b = a.vt_a;
  b.tt_b = a.tt_a;
  b.dpt_b = a.dpt_a;
  if (!lap1_next) hipLaunchKernelGGL(hv_dss_update_kernel<0>, gg, dim3(256), 0, st, b);
  else if (a.mixed) hipLaunchKernelGGL(hv_dss_update_kernel<2>, gg, dim3(256), 0, st, b);
  else hipLaunchKernelGGL(hv_dss_update_kernel<1>, gg, dim3(256), 0, st, b);
}

}  // namespace hx
