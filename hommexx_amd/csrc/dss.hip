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

__global__ __launch_bounds__(256) void dss_kernel(DssArgs a) {
  const int nblk = gridDim.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const size_t slab = (size_t)b * 16 + (threadIdx.x >> 4);
  const int pt = threadIdx.x & 15;
  const size_t nslab = (size_t)a.nelem * a.nlev;
  if (slab >= nslab) return;
  const int ie = (int)(slab / a.nlev), k = (int)(slab % a.nlev);
  const int *src = a.src + ((size_t)ie * 16 + pt) * 3;
  const int s0 = src[0], s1 = src[1], s2 = src[2];
  const double r = a.rsmp ? a.rsmp[(size_t)ie * 16 + pt] : 1.0;
  for (int f = 0; f < a.nf; ++f) {
    const DssField &F = a.f[f];
    const size_t off = (size_t)k * F.lstride;
    double x = F.in[(size_t)ie * F.estride + off + pt];
    if (s0 >= 0) x += F.in[(size_t)(s0 >> 4) * F.estride + off + (s0 & 15)];
    if (s1 >= 0) x += F.in[(size_t)(s1 >> 4) * F.estride + off + (s1 & 15)];
    if (s2 >= 0) x += F.in[(size_t)(s2 >> 4) * F.estride + off + (s2 & 15)];
    if (a.rsmp) x *= r;
    F.out[(size_t)ie * F.estride + off + pt] = x;
  }
}

void launch_dss(hipStream_t st, const DssArgs &a) {
  const size_t nslab = (size_t)a.nelem * a.nlev;
  const int blocks = (int)((nslab + 15) / 16);
  hipLaunchKernelGGL(dss_kernel, dim3(blocks), dim3(256), 0, st, a);
}

}  // namespace hx
