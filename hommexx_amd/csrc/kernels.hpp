// Kernel argument blocks and host-side launchers of the preqx step.
// Device layouts (all FP64, one array per time level):
//   scalar 3-D   [ie][lev][pt]            pt = igp*4 + jgp (Fortran (np,np,nlev))
//   vector 3-D   [ie][lev][comp][pt]      (Fortran (np,np,2,nlev))
//   tracers      [ie][q][lev][pt]         (Fortran (np,np,nlev,qsize))
//   2-D scalar   [ie][pt],  2-D tensor [ie][idim*2+jdim][pt]
#pragma once
#include <hip/hip_runtime.h>

#include "sphere32.hpp"

namespace hx {

struct CaarArgs {
  MetricPtrs m;
  int nelem, nlev, qsize;
  const double *v_n0, *t_n0, *dp_n0, *v_nm1, *t_nm1, *dp_nm1;
  const double *qdp_n0;  // moist: tracer block of the n0_qdp level (tracer 0 used), else nullptr
  double *vn0, *omega, *phi;
  double *divv;                 // scratch: div(v dp) per level
  int rsplit0;                  // rsplit == 0: Eulerian vertical advection
  const double *hybi;           // nlev+1 interface b coefficients (rsplit == 0)
  double *eta_acc;              // eta_dot_dpdn accumulator (rsplit == 0)
  double *out_v, *out_t, *out_dp;
  double dt, eta_ave_w, p_top;  // p_top = hyai(0)*ps0
};

constexpr int DSS_MAX_FIELDS = 8;
struct DssField {
  const double *in;
  double *out;
  long long estride;  // elements
  int lstride;        // levels
};
struct DssArgs {
  DssField f[DSS_MAX_FIELDS];
  int nf, nelem, nlev;
  int f32;               // fields hold FP32 values (accumulated in FP64)
  const int *src;        // [ie][16][3] global point ids (ie*16 + pt) or -1
  const double *rsmp;    // multiply by rspheremp if non-null
  // lap1 (FP64 state fields v0, v1, T, dp3d in f[0..3]): also the first
  // hyperviscosity laplacians of the assembled state, into lap_v, lap_t,
  // lap_dp (1: FP64, 2: the mixed form's FP32 laplacians)
  int lap1;
  MetricPtrs m;
  double *lap_v, *lap_t, *lap_dp;
  double nu_ratio1;
};

struct HvArgs {
  MetricPtrs m;
  const int *src;
  int nelem, nlev;
  double *v, *t, *dp;                // state at np1 (in/out)
  double *vt_a, *tt_a, *dpt_a;      // pre-exchange buffers
  double *vt_b, *tt_b, *dpt_b;      // second buffers
  double *dpdiss_ave, *dpdiss_bih;
  double dt, eta_ave_w, nu, nu_s, nu_p, nu_top, nu_ratio1, nu_ratio2;
  int nsub;
  int mixed;                        // FP32 laplacians, FP32 first-laplacian buffers
  int consthv;                      // 0: tensor form of the second laplacian
};

// The first laplacians of hyperviscosity at one (element, level, point) of
// a state (t, dp, v0, v1), into the pre-exchange buffers: FP64, or (F32) the
// mixed form's FP32 laplacians of the values less their slab reference,
// stored as FP32.  Every lane of the slab takes part (cross-lane reads).
template <bool F32>
__device__ __forceinline__ void hv_first_laplacians(const MetricPtrs &m, int ie, int pt, double nu_ratio1,
                                                    double t, double dp, double v0, double v1, double *lt,
                                                    double *ldp, double *lv, size_t s, size_t v) {
  if (F32) {
    ElemLaneF e;
    load_elemf(m, ie, pt, e);
    DvvLaneF d;
    load_dvvf(m.dvv, pt, d);
    ((float *)lt)[s] = sph_laplacef(e, d, (float)(t - slab_ref(t)));
    ((float *)ldp)[s] = sph_laplacef(e, d, (float)(dp - slab_ref(dp)));
    float l0, l1;
    sph_vlaplace_contraf(e, d, (float)nu_ratio1, (float)v0, (float)v1, l0, l1);
    ((float *)lv)[v] = l0;
    ((float *)lv)[v + 16] = l1;
  } else {
    ElemLane e;
    load_elem(m, ie, pt, e);
    DvvLane d;
    load_dvv(m.dvv, pt, d);
    lt[s] = sph_laplace(e, d, t);
    ldp[s] = sph_laplace(e, d, dp);
    double l0, l1;
    sph_vlaplace_contra(e, d, nu_ratio1, v0, v1, l0, l1);
    lv[v] = l0;
    lv[v + 16] = l1;
  }
}

// Layout of the tracer pre-exchange buffers (qdp_pre, qbih_pre, fdss_pre):
// levels fastest in chunks of 8 levels, [ie][field][lev/8][pt][lev%8].  One
// chunk of an element and field is 16 points x 8 levels (1 KB of FP64, 8
// lines of 128 B holding two GLL points each), so the block that gathers an
// element's 8-level group uses every line it fetches in full.
__host__ __device__ inline size_t lc_field(int nlev) { return (size_t)((nlev + 7) >> 3) * 128; }
__host__ __device__ inline size_t lc_index(int pt, int k) { return (size_t)(k >> 3) * 128 + pt * 8 + (k & 7); }

// Tracer bounds qlim: per element and tracer, each level's minimum and
// maximum side by side (one 16-byte load), [ie][q][lev][min,max], rows padded
// to a multiple of 8 levels (one 128-byte line per 8 levels).
__host__ __device__ inline size_t ql_row(int nlev) { return (size_t)((nlev + 7) >> 3) * 16; }
__host__ __device__ inline size_t ql_min(int k) { return (size_t)k * 2; }
constexpr int QL_MAX = 1;  // offset of a level's maximum from its minimum

struct EulerArgs {
  MetricPtrs m;
  const int *src, *nbr;
  int nelem, nlev, qsize;
  const double *qdp_n0;       // [ie][q][lev][pt]
  double *qdp_pre;            // pre-DSS output
  double *qdp_np1;            // final output (DSS)
  double *qbih, *qbih_pre;    // biharmonic buffers
  double *qlim, *qlim_new;    // ql_row / ql_min layout
  const double *derived_dp, *divdp, *divdp_proj, *vn0, *dpdiss_ave, *dpdiss_bih, *dp0;
  double *divdp_w, *divdp_proj_w;
  double *fdss, *fdss_pre;    // the extra DSS field and its pre-exchange copy
  double dt, rhs_multiplier, rhs_viss, nu_p, nu_q;
  int limiter_option;
  const double *qdp_avg_n0;   // stage 3: Qdp(n0_qdp) for the time average, else nullptr
  const double *qdp_pre_in, *fdss_pre_in;  // stage-1 pre-exchange buffers gathered by the fused stage 2
  int mixed;                  // FP32 flux divergence
  int bih32;                  // FP32 tracer biharmonic (qbih, qbih_pre hold floats)
  int consthv;                // 0: tensor laplacian in the biharmonic's second half
};

struct RemapArgs {
  int nelem, nlev, qsize, remap_alg;
  int nstates;                // 3: remap v, T and the tracers; 0 (rsplit == 0): tracers only
  const double *eta;          // rsplit == 0: eta_dot_dpdn, source = target + dt*(eta(k+1) - eta(k))
  double *src;                // rsplit == 0: scratch for the source thickness
  double dt;
  double *v, *t;              // state at np1
  const double *dp;           // source thickness dp3d(np1)
  double *ps;                 // ps_v(np1) [ie][pt]
  double *qdp;                // tracers at np1_qdp
  double *Q;                  // tracer mixing ratio out (update_q)
  const double *hyai, *hybi;  // nlev+1
  double ps0, hyai0;
  int *error;                 // set to 1 on invalid thickness, 2 on unsupported deformation
};

// Adjoint of one CAAR stage (dry).  The stage is
//   x_out = DSS_r( sph*(x_a + h*RHS(x_b)) );
// given lp = DSS_r(adjoint of x_out), the kernels add the adjoint of x_b into
// b (overwrite or accumulate) and c1*sph*lp into a1 (and c2*sph*lp into a2).
struct AdjCaarArgs {
  MetricPtrs m;
  int nelem, nlev;
  const double *v, *t, *dp;            // the stage's x_b
  const double *lp_v, *lp_t, *lp_dp;   // DSS_r of the output adjoint
  const double *hybi;                  // nlev+1
  double h, p_top;
  int rsplit0;                         // rsplit == 0: Eulerian vertical advection terms
  // scratch (one 3-D scalar each; stot is 2-D)
  double *div, *p, *c, *stot, *pb, *phib, *etab, *ombp, *divb, *lv, *lt, *ldp;
  double *b_v, *b_t, *b_dp;
  int b_acc;
  double *a1_v, *a1_t, *a1_dp;
  double c1;
  double *a2_v, *a2_t, *a2_dp;         // optional (a2_v == nullptr: none)
  double c2;
};

// Adjoint of one hyperviscosity cycle: x -> xo, adjoint l (in/out, in place)
struct AdjHvArgs {
  MetricPtrs m;
  int nelem, nlev;
  const double *x_v, *xo_v;            // velocity before and after the cycle
  double *l_v, *l_t, *l_dp;
  double *s_v, *s_t, *s_dp;            // DSS inputs
  const double *m_v, *m_t, *m_dp;      // DSS outputs
  double dt, nu, nu_s, nu_p, nu_top, nu_ratio1, nu_ratio2;
};

// Adjoint of the PPM remap of v and T (rsplit > 0) from the source
// thicknesses dp to the hybrid levels of ps = hyai(0) ps0 + sum dp
constexpr int REMAP_GRID = 5, REMAP_FIELD = 15;
struct AdjRemapArgs {
  int nelem, nlev, alg;
  const double *v, *t, *dp;            // remap input (the state after hyperviscosity)
  double *l_v, *l_t;                   // in: adjoint of the remapped v, T; out: of the input v, T
  double *l_dp;                        // out: adjoint of dp
  const double *l_ps;                  // adjoint of ps(np1)
  const double *hyai, *hybi;
  double ps0;
  double *scr;                         // adj_remap_scratch() doubles
};
size_t adj_remap_scratch(int nelem, int nlev);

// launchers (each enqueues on `st`)
void launch_adj_remap(hipStream_t st, const AdjRemapArgs &a);
void launch_adj_caar(hipStream_t st, const AdjCaarArgs &a);
void launch_adj_hv_pre(hipStream_t st, const AdjHvArgs &a);
void launch_adj_hv_a(hipStream_t st, const AdjHvArgs &a);
void launch_adj_hv_b(hipStream_t st, const AdjHvArgs &a);
void launch_adj_fill_dp(hipStream_t st, int nelem, int nlev, const double *lps, double *ldp);
void launch_adj_ps(hipStream_t st, int nelem, int nlev, const double *hybi, const double *ldp, double *lps);
void launch_adj_zero(hipStream_t st, size_t n, double *p);
void launch_dp3d_from_ps(hipStream_t st, int nelem, int nlev, const double *hyai_d, const double *hybi_d, double ps0,
                         const double *ps, double *dp);
void launch_prim_step_init(hipStream_t st, int nelem, int nlev, bool zero_diss, const double *dp_n0,
                           double *eta_dot, double *vn0, double *omega, double *dpa, double *dpb, double *derived_dp);
void launch_caar(hipStream_t st, const CaarArgs &a);
void launch_rk_combo(hipStream_t st, int nelem, int nlev, double *v_nm1, double *t_nm1, double *dp_nm1,
                     const double *v_n0, const double *t_n0, const double *dp_n0);
void launch_dss(hipStream_t st, const DssArgs &a);
// lap1_done: the first laplacians are already in vt_a, tt_a, dpt_a (formed by
// the previous subcycle's update); lap1_next: the update forms them for the
// next subcycle (FP64 form)
void launch_hv_cycle(hipStream_t st, const HvArgs &a, bool lap1_done = false, bool lap1_next = false);
void launch_euler_divdp_qmm(hipStream_t st, const EulerArgs &a);
void launch_euler_nbr_minmax(hipStream_t st, const EulerArgs &a);
void launch_euler_advect(hipStream_t st, const EulerArgs &a, int stage);
void launch_euler_advect_gather(hipStream_t st, const EulerArgs &a);
void launch_euler_dss(hipStream_t st, const EulerArgs &a, int avg);
// the stage-2 exchange with the stage-3 bounds and first laplacian of the
// biharmonic (a.rhs_multiplier 2); false: not launched (form not covered)
bool launch_euler_dss_qmm(hipStream_t st, const EulerArgs &a);
void launch_euler_qmm_lap(hipStream_t st, const EulerArgs &a);
void launch_euler_bih_post(hipStream_t st, const EulerArgs &a);
void launch_state_forcing(hipStream_t st, int nelem, int nlev, double dt, const double *ft, const double *fm,
                          double *t, double *v);
void launch_tracer_forcing(hipStream_t st, int nelem, int nlev, int qsize, bool moist, double dt, const double *fq,
                           double *qdp, double *Q, const double *hyai, const double *hybi, double ps0, double *ps);
// Diagnostics (Diagnostics.cpp:37-185): energy sums of dynamics level
// (v, t, ps) and, with use_cpstar, tracer 0 of qdp; tracer sums of qdp and Q.
struct DiagArgs {
  int nelem, nlev, qsize, use_cpstar;
  const double *v, *t, *ps, *phis, *qdp, *Q;
  const double *hyai, *hybi;  // nlev+1
  double ps0;
  double *out;  // energy: [4][nelem*16]; scalars: [2][nelem*qsize*16]
};
void launch_diag_energy(hipStream_t st, const DiagArgs &a);
void launch_diag_scalars(hipStream_t st, const DiagArgs &a);

// largest nlev of the remap kernel: an element's target table (10 B per
// level and column) and a grid chunk (20 KB) fit in 64 KB of LDS
constexpr int REMAP_MAX_NLEV = 256;
void launch_remap(hipStream_t st, const RemapArgs &a);
void launch_ps_from_dp(hipStream_t st, int nelem, int nlev, double hyai0, double ps0, const double *dp, double *ps);
void launch_zero_int(hipStream_t st, int *p);

// Streaming (non-temporal) store for outputs that only a later kernel reads:
// the line is not kept in L2, so the neighbour copies an exchange kernel
// gathers stay cached longer.
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

// Block barrier that orders LDS accesses only.  __syncthreads() also waits
// for every global load in flight; with this one, loads issued before the
// barrier (the next field's gathers) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// An index (or address) the compiler cannot see through: the selected source
// of a gather stays one unconditional load (otherwise the select of addresses
// is turned back into a branch around the neighbour load and a wait).
__device__ __forceinline__ size_t opaque_index(size_t i) {
  asm volatile("" : "+v"(i));
  return i;
}
template <typename P>
__device__ __forceinline__ P opaque_ptr(P p) {
  asm volatile("" : "+v"(p));
  return p;
}

// A pointer into global memory, kept in that address space: a generic
// pointer that went through opaque_ptr would be loaded through flat
// instructions, whose waits also cover the LDS operations.
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;

// DSS gather of one point: the own value r[0] and up to three neighbour
// copies r[1..3] from sources s0, s1, s2 (-1: absent), summed in that order.
// The loads are unconditional (an absent source re-reads the own value), so
// all four are in flight together instead of one branch and wait each; the
// absent terms are dropped by a select, which rounds exactly like the
// conditional sum.
template <typename T>
__device__ __forceinline__ double gather_sum(const T (&r)[4], int s0, int s1, int s2) {
  double x = r[0];
  x = s0 >= 0 ? x + (double)r[1] : x;
  x = s1 >= 0 ? x + (double)r[2] : x;
  x = s2 >= 0 ? x + (double)r[3] : x;
  return x;
}

// XCD-aware block order: consecutive logical blocks stay on one XCD (one L2)
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Exchange (gather) kernels walk one field at a time (blockIdx.y) through
// groups of GL = 8 levels, element by element: a 256-thread block is two
// elements x 8 levels, 16 lanes per (element, level).  The neighbour copies a
// block reads were loaded by blocks a few hundred positions earlier in the
// same level group (a few hundred KB), so they are still in the XCD's L2.
constexpr int GL = 8;
inline int group_blocks_x(int nelem, int nlev) {
  const long long ns = (long long)nelem * ((nlev + GL - 1) / GL) * GL;
  const int nb = (int)((ns + 15) / 16);
  return (nb + 7) & ~7;  // multiple of 8: blockIdx.x % 8 is then the XCD
}
__device__ __forceinline__ bool group_slab(int nelem, int nlev, int &ie, int &k, int &pt) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const long long s = (long long)lb * 16 + (threadIdx.x >> 4);
  pt = threadIdx.x & 15;
  const int kk = (int)(s & (GL - 1));
  const long long r = s >> 3;
  ie = (int)(r % nelem);
  const int g = (int)(r / nelem);
  k = g * GL + kk;
  return g < (nlev + GL - 1) / GL && k < nlev;
}

}  // namespace hx
