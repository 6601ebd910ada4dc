// Direct stiffness summation on one GPU (BoundaryExchange.cpp:285-618 in the
// reference, all connections LOCAL).  Instead of packing edge buffers, every
// output point gathers the neighbour copies it needs from the pre-exchange
// field.  The per-point source list (built on the host from the connection
// table) is already in the reference's accumulation order -- for k = 0..3
// the S, N, W, E edges, then the SW, SE, NW, NE corners -- so
//     out(p) = ((own(p) + c1) + c2) + c3     then  * rspheremp(p)
// rounds exactly like the reference's unpack.
#include "kernels.hpp"

namespace hx {

// One lane = one (element, level, point) for every field of the call: the
// source table is read once and all the fields' loads are issued before the
// first sum, so each lane has up to 4*nf independent loads in flight.
template <int LAP1>
__global__ __launch_bounds__(256) void dss_kernel(DssArgs a) {
  int ie, k, pt;
  if (!group_slab(a.nelem, a.nlev, ie, k, pt)) return;
  const int *src = a.src + ((size_t)ie * 16 + pt) * 3;
  const int s0 = src[0], s1 = src[1], s2 = src[2];
  const double r0 = a.rsmp ? a.rsmp[(size_t)ie * 16 + pt] : 1.0;
  // all loads of all fields are issued before the first sum (gather_sum);
  // an absent source re-reads the own value
  auto at = [&](const DssField &F, int src, size_t own) {
    return opaque_index(src >= 0 ? (size_t)(src >> 4) * F.estride + (size_t)k * F.lstride + (src & 15) : own);
  };
  if (!LAP1 && a.f32) {
    // FP32 fields of the mixed form, accumulated in FP64
    float r[DSS_MAX_FIELDS][4];
#pragma unroll
    for (int f = 0; f < DSS_MAX_FIELDS; ++f) {
      if (f < a.nf) {
        const DssField &F = a.f[f];
        const float *in = (const float *)F.in;
        const size_t own = (size_t)ie * F.estride + (size_t)k * F.lstride + pt;
        r[f][0] = in[own];
        r[f][1] = in[at(F, s0, own)];
        r[f][2] = in[at(F, s1, own)];
        r[f][3] = in[at(F, s2, own)];
      }
    }
#pragma unroll
    for (int f = 0; f < DSS_MAX_FIELDS; ++f) {
      if (f < a.nf) {
        const DssField &F = a.f[f];
        const double x = gather_sum(r[f], s0, s1, s2);
        ((float *)F.out)[(size_t)ie * F.estride + (size_t)k * F.lstride + pt] = (float)(a.rsmp ? x * r0 : x);
      }
    }
    return;
  }
  double r[DSS_MAX_FIELDS][4];
#pragma unroll
  for (int f = 0; f < DSS_MAX_FIELDS; ++f) {
    if (f < a.nf) {
      const DssField &F = a.f[f];
      const size_t own = (size_t)ie * F.estride + (size_t)k * F.lstride + pt;
      r[f][0] = F.in[own];
      r[f][1] = F.in[at(F, s0, own)];
      r[f][2] = F.in[at(F, s1, own)];
      r[f][3] = F.in[at(F, s2, own)];
    }
  }
  double y[DSS_MAX_FIELDS];
#pragma unroll
  for (int f = 0; f < DSS_MAX_FIELDS; ++f) {
    if (f < a.nf) {
      const DssField &F = a.f[f];
      const double x = gather_sum(r[f], s0, s1, s2);
      y[f] = a.rsmp ? x * r0 : x;
      F.out[(size_t)ie * F.estride + (size_t)k * F.lstride + pt] = y[f];
    }
  }
  if (LAP1) {
    // the first laplacians of hyperviscosity (the work of hv_lap1_kernel on
    // the stored state), from the assembled values in registers
    const size_t i = ((size_t)ie * a.nlev + k) * 16 + pt, v = ((size_t)ie * a.nlev + k) * 32 + pt;
    hv_first_laplacians<LAP1 == 2>(a.m, ie, pt, a.nu_ratio1, y[2], y[3], y[0], y[1], a.lap_t, a.lap_dp, a.lap_v, i, v);
  }
}

void launch_dss(hipStream_t st, const DssArgs &a) {
  if (a.nf <= 0) return;
  const dim3 g(group_blocks_x(a.nelem, a.nlev));
  if (a.lap1 == 2) hipLaunchKernelGGL(dss_kernel<2>, g, dim3(256), 0, st, a);
  else if (a.lap1) hipLaunchKernelGGL(dss_kernel<1>, g, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(dss_kernel<0>, g, dim3(256), 0, st, a);
}

}  // namespace hx
