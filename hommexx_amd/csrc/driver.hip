// Host side of libhommexx_amd.so: the reference's iso_c_binding C-ABI
// (cxx_f90_interface.cpp, mpi/mpi_cxx_f90_interface.cpp, Hommexx_Session.cpp,
// prim_driver.cpp) over a device-resident context, plus hommexx_hip_*
// extensions for the test/benchmark harness.  Fortran passes scalars by
// reference and arrays as pointers to the flat elem_* arrays; time-level
// indices enter 1-based and leave prim_run_subcycle_c 0-based, exactly as in
// the reference.  Errors that the reference turns into MPI_Abort abort the
// process too unless hommexx_hip_set_abort_on_error(0) was called, in which
// case they are recorded for hommexx_hip_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "../../include/hommexx_amd.h"

using namespace hx;

// the array behind scratch_pointer() (see there): 16 Mi doubles, 128 MiB per GPU
constexpr size_t ADJ_SCRATCH_DOUBLES = (size_t)16 << 20;
static __device__ __align__(256) double scratch[ADJ_SCRATCH_DOUBLES];

namespace {

struct Params {
  int remap_alg = 1, limiter_option = 8, rsplit = 1, qsplit = 1, time_step_type = 5, energy_fixer = 0;
  int qsize = 0, state_frequency = 1, hypervis_order = 2, hypervis_subcycle = 1, ftype = -1;
  double nu = 0, nu_p = 0, nu_q = 0, nu_s = 0, nu_div = 0, nu_top = 0, hypervis_scaling = 0;
  double nu_ratio1 = 1, nu_ratio2 = 1;
  bool prescribed_wind = false, moist = false, disable_diagnostics = true, use_cpstar = false, use_sl = false;
  bool set = false;
};

struct Timer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  int count = 0;
  double total_ms = 0;
};

struct Ctx {
  // configuration
  bool session = false;
  // the reference's build-time PLEV and QSIZE_D (Makefile PREQX_PLEV, QSIZE_D)
  int nlev_cfg = HOMMEXX_PREQX_PLEV, qsize_d_cfg = HOMMEXX_QSIZE_D;
  bool abort_on_error = true;
  int last_error = 0;
  std::string msg;
  hipStream_t st = nullptr;
  // connectivity (host)
  int conn_nelem = -1;
  bool conn_final = false;
  std::vector<int> conn_rl, conn_rp;  // [ie*8 + pos] -> remote lid / pos (reference numbering)
  // params / dims
  Params p;
  double dvv[16] = {0};
  bool dvv_set = false;
  int nelem = 0, nlev = 0, qsize = 0, qsize_d = 0;
  bool elems_set = false, states_set = false, hv_set = false, be_set = false;
  // 2-D metric (device)
  double *D = nullptr, *Dinv = nullptr, *metinv = nullptr, *tensorvisc = nullptr, *vec_sph2cart = nullptr;
  double *fcor = nullptr, *mp = nullptr, *spheremp = nullptr, *rspheremp = nullptr, *metdet = nullptr,
         *phis = nullptr, *dvv_d = nullptr;
  bool consthv = true;
  int *dss_src = nullptr, *nbr = nullptr;
  // vertical coordinate
  double ps0 = 0, hyai0 = 0;
  std::vector<double> hyai, hybi;
  double *hyai_d = nullptr, *hybi_d = nullptr, *dp0_d = nullptr;
  // states
  double *v[3] = {}, *t[3] = {}, *dp[3] = {}, *ps[3] = {}, *qdp[2] = {}, *Q = nullptr;
  std::vector<double> qdp_tail;  // Qdp of tracers qsize..qsize_d-1 as received (never evolved)
  std::vector<double> fq_tail;   // FQ of tracers qsize..qsize_d-1 as received (never used)
  // derived
  double *phi = nullptr, *omega = nullptr, *eta_dot = nullptr, *dpdiss_ave = nullptr, *dpdiss_bih = nullptr,
         *derived_dp = nullptr, *divdp = nullptr, *divdp_proj = nullptr, *vn0 = nullptr;
  // forcing
  double *ft = nullptr, *fm = nullptr, *fq = nullptr;
  // scratch
  double *cv = nullptr, *ct = nullptr, *cdp = nullptr, *cdiv = nullptr;
  double *vt_a = nullptr, *tt_a = nullptr, *dpt_a = nullptr, *vt_b = nullptr, *tt_b = nullptr, *dpt_b = nullptr;
  double *qdp_pre = nullptr, *qbih = nullptr, *qbih_pre = nullptr, *qlim[2] = {}, *fdss_pre = nullptr;
  // stage-2 pre-exchange buffers: qdp_pre2 (later qbih) and an alias of ct,
  // which is free from the stage-2 advection until its exchange has consumed it
  double *qdp_pre2 = nullptr, *fdss_pre2 = nullptr;
  int *err_d = nullptr;
  int *err_h = nullptr;
  // diagnostics: the Fortran arrays of init_diagnostics_c (q, qvar, qmass,
  // q1mass, iener, iener_wet, kener, pener) and a device staging buffer
  double *dg[8] = {};
  bool dg_set = false;
  double *dg_dev = nullptr;
  // time levels (0-based)
  int nm1 = 0, n0 = 1, np1 = 2, nstep = 0, nstep0 = 0, n0_qdp = 0, np1_qdp = 1;
  // 0: FP64 step; 1: mixed form (FP32 sphere operators in hyperviscosity and
  // the tracer flux divergence / biharmonic)
  int form = 0;
  // debugging: stop prim_run_subcycle_c after a phase (0: never)
  int debug_stop = 0;
  bool stopped = false;
  // profiling
  bool prof = false;
  std::string prof_only;  // non-empty: time only the phase of this name
  std::map<std::string, Timer> timers;
  std::vector<void *> allocs;
};

Ctx *g = nullptr;

Ctx &C() {
  if (!g) g = new Ctx();
  return *g;
}

void fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  Ctx &c = C();
  c.last_error = code;
  c.msg = buf;
  if (c.abort_on_error) {
    fprintf(stderr, "hommexx_amd error %d: %s\n", code, buf);
    fflush(stderr);
    abort();
  }
}

bool hipok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fail(100, "%s: %s", what, hipGetErrorString(e));
    return false;
  }
  return true;
}

template <typename T>
T *dalloc(size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  if (!hipok(hipMalloc(&p, n * sizeof(T)), "hipMalloc")) return nullptr;
  // hipMemset runs on the null stream, which the library's non-blocking
  // stream does not wait for: finish it before any kernel can touch p
  hipMemset(p, 0, n * sizeof(T));
  if (!hipok(hipDeviceSynchronize(), "hipMemset")) return nullptr;
  C().allocs.push_back(p);
  return (T *)p;
}

void h2d(void *d, const void *h, size_t bytes) { hipok(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice), "h2d"); }
void d2h(void *h, const void *d, size_t bytes) { hipok(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost), "d2h"); }

void adj_reset();

void free_all() {
  Ctx &c = C();
  adj_reset();
  if (c.st) hipStreamSynchronize(c.st);
  for (void *p : c.allocs) hipFree(p);
  c.allocs.clear();
  for (auto &kv : c.timers)
    for (auto &e : kv.second.pending) {
      hipEventDestroy(e.first);
      hipEventDestroy(e.second);
    }
  c.timers.clear();
  if (c.err_h) hipHostFree(c.err_h);
  c.err_h = nullptr;
  c.dg_dev = nullptr;
}

// ---------------------------------------------------------------- timing
struct Scope {
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  explicit Scope(const char *n) : name(n) {
    Ctx &c = C();
    if (!c.prof || (!c.prof_only.empty() && c.prof_only != n)) return;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, c.st);
  }
  ~Scope() {
    Ctx &c = C();
    if (!c.prof || !a) return;
    hipEventRecord(b, c.st);
    c.timers[name].pending.push_back({a, b});
  }
};

void flush_timers() {
  Ctx &c = C();
  hipStreamSynchronize(c.st);
  for (auto &kv : c.timers) {
    for (auto &e : kv.second.pending) {
      float ms = 0;
      hipEventElapsedTime(&ms, e.first, e.second);
      kv.second.total_ms += ms;
      kv.second.count += 1;
      hipEventDestroy(e.first);
      hipEventDestroy(e.second);
    }
    kv.second.pending.clear();
  }
}

// ------------------------------------------------------ connectivity
// point lists of the 8 connections (ConnectivityHelpers.hpp), pt = igp*4+jgp
const int EDGE_FWD[4][4] = {{0, 1, 2, 3}, {12, 13, 14, 15}, {0, 4, 8, 12}, {3, 7, 11, 15}};
const int CORNER_PT[4] = {0, 3, 12, 15};
const int EDGE_DIR[4][4] = {{1, 0, 0, 1}, {0, 1, 1, 0}, {0, 1, 1, 0}, {1, 0, 0, 1}};

void build_dss_tables() {
  Ctx &c = C();
  const int ne = c.nelem;
  // slot (receiver elem, receiver pos)[k] <- (sender elem, sender point)
  std::vector<int> slot((size_t)ne * 8 * 4, -1);
  for (int a = 0; a < ne; ++a)
    for (int pa = 0; pa < 8; ++pa) {
      const int b = c.conn_rl[a * 8 + pa];
      if (b < 0) continue;
      const int pb = c.conn_rp[a * 8 + pa];
      if (pa < 4) {
        const int dir = EDGE_DIR[pa][pb];
        for (int k = 0; k < 4; ++k) {
          const int spt = EDGE_FWD[pa][dir ? 3 - k : k];
          slot[((size_t)b * 8 + pb) * 4 + k] = a * 16 + spt;
        }
      } else {
        slot[((size_t)b * 8 + pb) * 4 + 0] = a * 16 + CORNER_PT[pa - 4];
      }
    }
  std::vector<int> src((size_t)ne * 16 * 3, -1), cnt((size_t)ne * 16, 0);
  std::vector<int> nbr((size_t)ne * 8, -1);
  auto push = [&](int e, int pt, int s) {
    const size_t i = (size_t)e * 16 + pt;
    if (cnt[i] >= 3) {
      fail(10, "more than three neighbour copies at element %d point %d", e, pt);
      return;
    }
    src[i * 3 + cnt[i]++] = s;
  };
  for (int e = 0; e < ne; ++e) {
    for (int k = 0; k < 4; ++k)
      for (int ed = 0; ed < 4; ++ed) {
        const int s = slot[((size_t)e * 8 + ed) * 4 + k];
        if (s >= 0) push(e, EDGE_FWD[ed][k], s);
      }
    for (int cr = 0; cr < 4; ++cr) {
      const int s = slot[((size_t)e * 8 + 4 + cr) * 4];
      if (s >= 0) push(e, CORNER_PT[cr], s);
    }
    for (int pos = 0; pos < 8; ++pos) {
      const int s = slot[((size_t)e * 8 + pos) * 4];
      nbr[(size_t)e * 8 + pos] = s >= 0 ? s / 16 : -1;
    }
  }
  c.dss_src = dalloc<int>(src.size());
  c.nbr = dalloc<int>(nbr.size());
  h2d(c.dss_src, src.data(), src.size() * sizeof(int));
  h2d(c.nbr, nbr.data(), nbr.size() * sizeof(int));
}

MetricPtrs metric() {
  Ctx &c = C();
  MetricPtrs m;
  m.D = c.D;
  m.Dinv = c.Dinv;
  m.metinv = c.metinv;
  m.fcor = c.fcor;
  m.mp = c.mp;
  m.spheremp = c.spheremp;
  m.rspheremp = c.rspheremp;
  m.metdet = c.metdet;
  m.phis = c.phis;
  m.dvv = c.dvv_d;
  m.tensorvisc = c.tensorvisc;
  m.vec_sph2cart = c.vec_sph2cart;
  return m;
}

bool ready(const char *who) {
  Ctx &c = C();
  if (!(c.p.set && c.dvv_set && c.elems_set && c.states_set && c.hv_set && c.conn_final)) {
    fail(20, "%s: the library is not fully initialised (params %d, derivative %d, elements %d, states %d, hvcoord %d, "
             "connectivity %d)",
         who, c.p.set, c.dvv_set, c.elems_set, c.states_set, c.hv_set, c.conn_final);
    return false;
  }
  if (c.conn_nelem != c.nelem) {
    fail(21, "%s: connectivity has %d elements but the element data has %d", who, c.conn_nelem, c.nelem);
    return false;
  }
  return true;
}

void alloc_work() {
  Ctx &c = C();
  if (c.be_set) return;
  const size_t e3 = (size_t)c.nelem * c.nlev * 16;
  const size_t q3 = e3 * (c.qsize ? c.qsize : 1);
  // pre-exchange buffers: whole 8-level chunks per element and field
  const size_t e3c = (size_t)c.nelem * lc_field(c.nlev), q3c = e3c * (c.qsize ? c.qsize : 1);
  c.cv = dalloc<double>(2 * e3);
  c.ct = dalloc<double>(e3c > e3 ? e3c : e3);  // also the stage-2 fdss_pre
  c.cdp = dalloc<double>(e3);
  c.cdiv = dalloc<double>(e3);
  c.vt_a = dalloc<double>(2 * e3);
  c.vt_b = dalloc<double>(2 * e3);
  c.tt_a = dalloc<double>(e3);
  c.tt_b = dalloc<double>(e3);
  c.dpt_a = dalloc<double>(e3);
  c.dpt_b = dalloc<double>(e3);
  // Five 40-F3 tracer uses share two allocations (at Q = 40); no kernel
  // writes the allocation it gathers from:
  //   qdp_pre : stage-1 pre-exchange (-> the stage-2 gather), then the first
  //             laplacian of the biharmonic (the stage-2 exchange -> bih_post),
  //             then the stage-3 pre-exchange (-> euler_dss)
  //   qdp_pre2: stage-2 pre-exchange (-> the stage-2 exchange), then the
  //             biharmonic result qbih (bih_post -> stage-3 advection)
  c.qdp_pre = dalloc<double>(q3c);
  c.qbih_pre = c.qdp_pre;
  c.qdp_pre2 = dalloc<double>(q3c > q3 ? q3c : q3);
  c.qbih = c.qdp_pre2;
  c.qlim[0] = dalloc<double>((size_t)c.nelem * (c.qsize ? c.qsize : 1) * ql_row(c.nlev));
  c.qlim[1] = dalloc<double>((size_t)c.nelem * (c.qsize ? c.qsize : 1) * ql_row(c.nlev));
  c.fdss_pre = dalloc<double>(e3c);
  c.fdss_pre2 = c.ct;
  c.err_d = dalloc<int>(1);
  if (!c.err_h) hipHostMalloc((void **)&c.err_h, sizeof(int));
  c.be_set = true;
}

// ------------------------------------------------------ step phases
// one prognostic state: v [ie][lev][2][pt], T and dp3d [ie][lev][pt]
struct StateBuf {
  double *v, *t, *dp;
};

// out = DSS_r( sph*(a + dt*RHS(b)) ): one CAAR stage with its DSS.  lap1:
// the DSS also forms the first hyperviscosity laplacians of out (FP64 form)
void caar_run(const StateBuf &sa, const StateBuf &sb, const StateBuf &out, double dt, double eta_ave_w,
              bool lap1 = false) {
  Ctx &c = C();
  CaarArgs a{};
  a.m = metric();
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.qsize = c.qsize;
  a.v_n0 = sb.v;
  a.t_n0 = sb.t;
  a.dp_n0 = sb.dp;
  a.v_nm1 = sa.v;
  a.t_nm1 = sa.t;
  a.dp_nm1 = sa.dp;
  a.qdp_n0 = (c.p.moist && c.qsize > 0) ? c.qdp[c.n0_qdp] : nullptr;
  a.vn0 = c.vn0;
  a.omega = c.omega;
  a.phi = c.phi;
  a.divv = c.cdiv;
  a.rsplit0 = c.p.rsplit == 0;
  a.hybi = c.hybi_d;
  a.eta_acc = c.eta_dot;
  a.out_v = c.cv;
  a.out_t = c.ct;
  a.out_dp = c.cdp;
  a.dt = dt;
  a.eta_ave_w = eta_ave_w;
  a.p_top = c.hyai0 * c.ps0;
  {
    Scope s("caar");
    launch_caar(c.st, a);
  }
  DssArgs x{};
  x.nf = 4;
  x.nelem = c.nelem;
  x.nlev = c.nlev;
  x.src = c.dss_src;
  x.rsmp = c.rspheremp;
  const long long es = (long long)c.nlev * 16, ev = (long long)c.nlev * 32;
  x.f[0] = {c.cv, out.v, ev, 32};
  x.f[1] = {c.cv + 16, out.v + 16, ev, 32};
  x.f[2] = {c.ct, out.t, es, 16};
  x.f[3] = {c.cdp, out.dp, es, 16};
  if (lap1) {
    x.lap1 = c.form == 1 && c.consthv ? 2 : 1;  // the mixed form's FP32 first laplacians
    x.m = metric();
    x.lap_v = c.vt_a;
    x.lap_t = c.tt_a;
    x.lap_dp = c.dpt_a;
    x.nu_ratio1 = c.p.nu_ratio1;
  }
  Scope s("caar_dss");
  launch_dss(c.st, x);
}

StateBuf level(int tl) {
  Ctx &c = C();
  return {c.v[tl], c.t[tl], c.dp[tl]};
}

void caar_stage(int nm1, int n0, int np1, double dt, double eta_ave_w, bool lap1 = false) {
  caar_run(level(nm1), level(n0), level(np1), dt, eta_ave_w, lap1);
}

// hyperviscosity's first laplacians are formed by the last RK stage's DSS
// (with at least one subcycle)
bool hv_lap1_fused() {
  Ctx &c = C();
  return c.p.hypervis_subcycle > 0;
}

bool stop_here(int phase) {
  Ctx &c = C();
  if (c.debug_stop == phase) c.stopped = true;
  return c.stopped;
}

void u3_5stage(int nm1, int n0, int np1, double dt, double eta_ave_w) {
  Ctx &c = C();
  caar_stage(n0, n0, nm1, dt / 5.0, eta_ave_w / 4.0);
  if (stop_here(1)) return;
  caar_stage(n0, nm1, np1, dt / 5.0, 0.0);
  caar_stage(n0, np1, np1, dt / 3.0, 0.0);
  caar_stage(n0, np1, np1, 2.0 * dt / 3.0, 0.0);
  {
    Scope s("rk_combo");
    launch_rk_combo(c.st, c.nelem, c.nlev, c.v[nm1], c.t[nm1], c.dp[nm1], c.v[n0], c.t[n0], c.dp[n0]);
  }
  caar_stage(nm1, np1, np1, 3.0 * dt / 4.0, 3.0 * eta_ave_w / 4.0, hv_lap1_fused());
}

HvArgs hv_args(const StateBuf &x, double dt, double eta_ave_w) {
  Ctx &c = C();
  HvArgs a{};
  a.m = metric();
  a.src = c.dss_src;
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.v = x.v;
  a.t = x.t;
  a.dp = x.dp;
  a.vt_a = c.vt_a;
  a.tt_a = c.tt_a;
  a.dpt_a = c.dpt_a;
  a.vt_b = c.vt_b;
  a.tt_b = c.tt_b;
  a.dpt_b = c.dpt_b;
  a.dpdiss_ave = c.dpdiss_ave;
  a.dpdiss_bih = c.dpdiss_bih;
  a.dt = dt / c.p.hypervis_subcycle;
  a.eta_ave_w = eta_ave_w;
  a.nu = c.p.nu;
  a.nu_s = c.p.nu_s;
  a.nu_p = c.p.nu_p;
  a.nu_top = c.p.nu_top;
  a.nu_ratio1 = c.p.nu_ratio1;
  a.nu_ratio2 = c.p.nu_ratio2;
  a.nsub = c.p.hypervis_subcycle;
  // the mixed form keeps the tensor form of hyperviscosity in FP64
  a.consthv = c.consthv;
  a.mixed = c.form == 1 && c.consthv;
  return a;
}

void hyperviscosity(int np1, double dt, double eta_ave_w) {
  Ctx &c = C();
  const HvArgs a = hv_args(level(np1), dt, eta_ave_w);
  const int nsub = c.p.hypervis_subcycle;
  const bool first = hv_lap1_fused();  // the first laplacians came with the last RK stage
  for (int cyc = 0; cyc < nsub; ++cyc) {
    Scope s("hv_cycle");
    launch_hv_cycle(c.st, a, cyc > 0 || first, cyc + 1 < nsub);
  }
}

void update_tracers_levels() {
  Ctx &c = C();
  const int it = c.nstep / c.p.qsplit;
  c.n0_qdp = it % 2 == 0 ? 0 : 1;
  c.np1_qdp = 1 - c.n0_qdp;
}

void leapfrog() {
  Ctx &c = C();
  const int tmp = c.np1;
  c.np1 = c.nm1;
  c.nm1 = c.n0;
  c.n0 = tmp;
  ++c.nstep;
}

void advec_tracers(double dt) {
  Ctx &c = C();
  update_tracers_levels();
  EulerArgs a{};
  a.m = metric();
  a.src = c.dss_src;
  a.nbr = c.nbr;
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.qsize = c.qsize;
  a.qdp_pre = c.qdp_pre;
  a.qbih = c.qbih;
  a.qbih_pre = c.qbih_pre;
  a.derived_dp = c.derived_dp;
  a.divdp = c.divdp;
  a.divdp_proj = c.divdp_proj;
  a.divdp_w = c.divdp;
  a.divdp_proj_w = c.divdp_proj;
  a.vn0 = c.vn0;
  a.dpdiss_ave = c.dpdiss_ave;
  a.dpdiss_bih = c.dpdiss_bih;
  a.dp0 = c.dp0_d;
  a.fdss_pre = c.fdss_pre;
  a.nu_p = c.p.nu_p;
  a.nu_q = c.p.nu_q;
  a.limiter_option = c.p.limiter_option;
  a.mixed = c.form == 1;
  a.bih32 = c.form == 1 && c.consthv;
  a.consthv = c.consthv;
  a.rhs_viss = 0.0;
  const double dts = dt / 2.0;
  a.dt = dts;
  int cur = 0;
  // stage 1 bounds: divdp, then qmin/qmax of Qdp(n0_qdp) (rhs_multiplier 0)
  a.rhs_multiplier = 0.0;
  a.qdp_n0 = c.qdp[c.n0_qdp];
  a.qlim = c.qlim[cur];
  {
    Scope s("euler_divdp_qmm");
    launch_euler_divdp_qmm(c.st, a);
  }
  auto nbr_minmax = [&]() {
    a.qlim = c.qlim[cur];
    a.qlim_new = c.qlim[1 - cur];
    Scope s("euler_nbr_minmax");
    launch_euler_nbr_minmax(c.st, a);
    cur = 1 - cur;
    a.qlim = c.qlim[cur];
  };
  nbr_minmax();
  bool qmm_done = false;  // stage-3 bounds and first laplacian formed by the stage-2 exchange
  for (int stage = 0; stage < 3; ++stage) {
    const int n0q = stage == 0 ? c.n0_qdp : c.np1_qdp;
    a.qdp_n0 = c.qdp[n0q];
    a.qdp_np1 = c.qdp[c.np1_qdp];
    a.rhs_multiplier = (double)stage;
    a.rhs_viss = stage == 2 ? 3.0 : 0.0;
    a.fdss = stage == 0 ? c.divdp_proj : stage == 1 ? c.eta_dot : c.omega;
    // (stage 3 reads qbih from the other allocation)
    a.qdp_pre = c.qdp_pre;
    a.qdp_avg_n0 = stage == 2 ? c.qdp[c.n0_qdp] : nullptr;
    // stage 2 gathers stage 1's exchange itself (the assembled stage-1 Qdp
    // is never stored) unless the caller stops after stage 1
    const bool fuse1 = c.debug_stop != 4;
    if (stage == 1 && fuse1) {
      EulerArgs b = a;
      b.qdp_pre_in = c.qdp_pre;
      b.fdss_pre_in = c.fdss_pre;
      b.qdp_pre = c.qdp_pre2;
      b.fdss_pre = c.fdss_pre2;
      {
        Scope s("euler_advect");
        launch_euler_advect_gather(c.st, b);
      }
      b.fdss = c.eta_dot;
      // the exchange also forms the stage-3 bounds and the first laplacian
      // of the biharmonic (the qmm_lap step below) when the form allows
      b.rhs_multiplier = 2.0;
      Scope s("euler_dss_qmm");
      qmm_done = launch_euler_dss_qmm(c.st, b);
      if (!qmm_done) {
        b.rhs_multiplier = 1.0;
        launch_euler_dss(c.st, b, 0);
      }
    } else {
      {
        Scope s("euler_advect");
        launch_euler_advect(c.st, a, stage + 1);
      }
      if (!(stage == 0 && fuse1)) {
        Scope s("euler_dss");
        launch_euler_dss(c.st, a, stage == 2);
      }
    }
    if (stage == 0 && stop_here(4)) return;
    if (stage == 1) {
      // stage-3 bounds and tracer biharmonic from the assembled stage-2 Qdp
      EulerArgs b = a;
      b.qdp_n0 = c.qdp[c.np1_qdp];
      b.rhs_multiplier = 2.0;
      b.rhs_viss = 3.0;
      if (!qmm_done) {
        Scope s("euler_qmm_lap");
        launch_euler_qmm_lap(c.st, b);
      }
      {
        Scope s("euler_bih_post");
        launch_euler_bih_post(c.st, b);
      }
      nbr_minmax();
    }
  }
}

void prim_step(double dt) {
  Ctx &c = C();
  {
    Scope s("prim_step_init");
    launch_prim_step_init(c.st, c.nelem, c.nlev, c.p.nu_p > 0, c.dp[c.n0], c.eta_dot, c.vn0, c.omega, c.dpdiss_ave,
                          c.dpdiss_bih, c.derived_dp);
  }
  for (int n = 0; n < c.p.qsplit; ++n) {
    if (n > 0) leapfrog();
    if (c.p.moist) update_tracers_levels();
    const double eta_ave_w = 1.0 / c.p.qsplit;
    u3_5stage(c.nm1, c.n0, c.np1, dt, eta_ave_w);
    if (c.stopped || stop_here(2)) return;
    hyperviscosity(c.np1, dt, eta_ave_w);
    if (stop_here(3)) return;
  }
  if (c.qsize > 0) advec_tracers(dt * c.p.qsplit);
  stop_here(5);
}

void remap(double dt_remap) {
  Ctx &c = C();
  RemapArgs a{};
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.qsize = c.qsize;
  a.nstates = c.p.rsplit > 0 ? 3 : 0;
  a.eta = c.eta_dot;
  a.src = c.cdiv;  // CAAR's scratch is free at the remap
  a.dt = dt_remap;
  a.remap_alg = c.p.remap_alg;
  a.v = c.v[c.np1];
  a.t = c.t[c.np1];
  a.dp = c.dp[c.np1];
  a.ps = c.ps[c.np1];
  a.qdp = c.qsize ? c.qdp[c.np1_qdp] : nullptr;
  a.Q = c.Q;
  a.hyai = c.hyai_d;
  a.hybi = c.hybi_d;
  a.ps0 = c.ps0;
  a.hyai0 = c.hyai0;
  a.error = c.err_d;
  Scope s("remap");
  launch_remap(c.st, a);
}

// ------------------------------------------------------------ adjoint
// Reverse-mode derivative of one dry rsplit = 0 step (no tracers on the
// dynamics path).  The forward stage states are recomputed into private
// buffers first; the step's model state is not touched.
// The stage states of one dynamics step: CAAR stage inputs U0..U4, the
// combination w = (5 u1 - u0)/4 and the HV cycle states H0..Hnsub (U0 is the
// step's input, H[nsub] its output)
struct Stages {
  std::vector<StateBuf> U, H;
  StateBuf W{};
};

struct AdjBufs {
  size_t f3 = 0;  // doubles in one 3-D scalar
  int nsub = 0;
  int nd = 0;  // dynamics steps per prim_run_subcycle_c
  Stages cur;                  // the stage states of one dynamics step
  std::vector<StateBuf> X;     // input states of the nd dynamics steps
  StateBuf LX{}, L0{}, L1{}, LP{}, S{}, M{};
  double *div, *p, *cc, *stot, *pb, *phib, *etab, *ombp, *divb, *lv, *lt, *ldp;
  double *rscr = nullptr;  // remap adjoint scratch (rsplit > 0)
  double *base = nullptr;
};
AdjBufs *adjb = nullptr;

// hommexx_hip_adjoint_segment: the stage states of every dynamics step of a
// segment of calls, so the reverse sweep recomputes none of them
struct SegBufs {
  size_t f3 = 0;
  int nsub = 0, nsets = 0;
  std::vector<Stages> sets;
  StateBuf x{};  // the input of the next call (forward pass)
  double *ps = nullptr;
};
SegBufs *segb = nullptr;

void adj_reset() {
  delete adjb;
  adjb = nullptr;
  delete segb;
  segb = nullptr;
}

StateBuf carve_state(double *&q, size_t f3) {
  StateBuf s{q, q + 2 * f3, q + 3 * f3};
  q += 4 * f3;
  return s;
}

// Fixed device scratch of hommexx_hip_adjoint_step_noalloc: ADJ_SCRATCH_DOUBLES
// (128 MiB) per GPU, held until the process ends.  In host code `scratch`
// only names the symbol: its device address comes from hipGetSymbolAddress,
// once per device with that device current, and only that address goes to
// kernels and copies.
std::map<int, double *> scratch_ptrs;  // device -> the address of its copy of scratch

bool scratch_setup() {
  int dev = 0;
  if (!hipok(hipGetDevice(&dev), "hipGetDevice")) return false;
  if (scratch_ptrs.count(dev)) return true;
  void *p = nullptr;
  if (!hipok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return false;
  scratch_ptrs[dev] = (double *)p;
  return true;
}

double *scratch_pointer() {
  int dev = 0;
  if (!hipok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
  if (!scratch_ptrs.count(dev) && !scratch_setup()) return nullptr;
  return scratch_ptrs[dev];
}

// doubles of the adjoint step's buffers for nd dynamics steps per call
size_t adj_total_doubles(int nd) {
  Ctx &c = C();
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  const int nstates = 5 + 1 + (c.p.hypervis_subcycle + 1) + 6 + nd;  // U, W, H, LX L0 L1 LP S M, X
  return (size_t)nstates * 4 * f3 + 12 * f3 + (size_t)c.nelem * 16 + adj_remap_scratch(c.nelem, c.nlev);
}

// carve the adjoint step's buffers from base (adj_total_doubles(nd) doubles)
void adj_layout(AdjBufs &b, double *base, int nd) {
  Ctx &c = C();
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  const int nsub = c.p.hypervis_subcycle;
  b.base = base;
  b.f3 = f3;
  b.nsub = nsub;
  b.nd = nd;
  double *q = b.base;
  b.cur.U.clear();
  b.cur.H.clear();
  b.X.clear();
  for (int i = 0; i < nd; ++i) b.X.push_back(carve_state(q, f3));
  for (int i = 0; i < 5; ++i) b.cur.U.push_back(carve_state(q, f3));
  b.cur.W = carve_state(q, f3);
  for (int i = 0; i <= nsub; ++i) b.cur.H.push_back(carve_state(q, f3));
  b.LX = carve_state(q, f3);
  b.L0 = carve_state(q, f3);
  b.L1 = carve_state(q, f3);
  b.LP = carve_state(q, f3);
  b.S = carve_state(q, f3);
  b.M = carve_state(q, f3);
  double **sc[] = {&b.div, &b.p, &b.cc, &b.pb, &b.phib, &b.etab, &b.ombp, &b.divb, &b.lt, &b.ldp};
  for (double **x : sc) {
    *x = q;
    q += f3;
  }
  b.lv = q;
  q += 2 * f3;
  b.stot = q;
  q += (size_t)c.nelem * 16;
  b.rscr = q;
}

bool adj_alloc(int nd) {
  Ctx &c = C();
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  if (adjb && adjb->f3 == f3 && adjb->nsub == c.p.hypervis_subcycle && adjb->nd == nd) return true;
  if (!adjb) adjb = new AdjBufs();
  double *base = dalloc<double>(adj_total_doubles(nd));  // owned by the session (freed with it)
  if (!base) return false;
  adj_layout(*adjb, base, nd);
  return true;
}

void copy_state(const StateBuf &dst, const StateBuf &src, size_t f3) {
  Ctx &c = C();
  hipMemcpyAsync(dst.v, src.v, 4 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
}

// y = DSS_r(x) for the four fields of a state
void dss_state(const StateBuf &x, const StateBuf &y) {
  Ctx &c = C();
  DssArgs a{};
  a.nf = 4;
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.src = c.dss_src;
  a.rsmp = c.rspheremp;
  const long long es = (long long)c.nlev * 16, ev = (long long)c.nlev * 32;
  a.f[0] = {x.v, y.v, ev, 32};
  a.f[1] = {x.v + 16, y.v + 16, ev, 32};
  a.f[2] = {x.t, y.t, es, 16};
  a.f[3] = {x.dp, y.dp, es, 16};
  launch_dss(c.st, a);
}

// adjoint of out = DSS_r(sph*(a + h*RHS(b))), lam = adjoint of out
// (b: the step's buffers, scratch included)
void caar_adjoint(AdjBufs &b, const StateBuf &xb, const StateBuf &lam, double h, const StateBuf &lb, bool b_acc,
                  const StateBuf &la1, double c1, const StateBuf *la2, double c2) {
  Ctx &c = C();
  dss_state(lam, b.LP);
  AdjCaarArgs a{};
  a.m = metric();
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.v = xb.v;
  a.t = xb.t;
  a.dp = xb.dp;
  a.lp_v = b.LP.v;
  a.lp_t = b.LP.t;
  a.lp_dp = b.LP.dp;
  a.hybi = c.hybi_d;
  a.h = h;
  a.p_top = c.hyai0 * c.ps0;
  a.rsplit0 = c.p.rsplit == 0;
  a.div = b.div;
  a.p = b.p;
  a.c = b.cc;
  a.stot = b.stot;
  a.pb = b.pb;
  a.phib = b.phib;
  a.etab = b.etab;
  a.ombp = b.ombp;
  a.divb = b.divb;
  a.lv = b.lv;
  a.lt = b.lt;
  a.ldp = b.ldp;
  a.b_v = lb.v;
  a.b_t = lb.t;
  a.b_dp = lb.dp;
  a.b_acc = b_acc;
  a.a1_v = la1.v;
  a.a1_t = la1.t;
  a.a1_dp = la1.dp;
  a.c1 = c1;
  if (la2) {
    a.a2_v = la2->v;
    a.a2_t = la2->t;
    a.a2_dp = la2->dp;
    a.c2 = c2;
  }
  launch_adj_caar(c.st, a);
}

// adjoint of one hyperviscosity cycle x -> xo; lam (in place)
void hv_adjoint(AdjBufs &b, const StateBuf &x, const StateBuf &xo, const StateBuf &lam, double dt_cycle) {
  Ctx &c = C();
  AdjHvArgs a{};
  a.m = metric();
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.x_v = x.v;
  a.xo_v = xo.v;
  a.l_v = lam.v;
  a.l_t = lam.t;
  a.l_dp = lam.dp;
  a.s_v = b.S.v;
  a.s_t = b.S.t;
  a.s_dp = b.S.dp;
  a.m_v = b.M.v;
  a.m_t = b.M.t;
  a.m_dp = b.M.dp;
  a.dt = dt_cycle;
  a.nu = c.p.nu;
  a.nu_s = c.p.nu_s;
  a.nu_p = c.p.nu_p;
  a.nu_top = c.p.nu_top;
  a.nu_ratio1 = c.p.nu_ratio1;
  a.nu_ratio2 = c.p.nu_ratio2;
  launch_adj_hv_pre(c.st, a);
  dss_state(b.S, b.M);
  launch_adj_hv_a(c.st, a);
  dss_state(b.S, b.M);
  launch_adj_hv_b(c.st, a);
}

// Fortran (np,np,nlev,...) blocks are already [lev][pt] per element; only the
// time-level / tracer strides differ between the Fortran arrays and ours.
void pull_levels(double *dev, const double *f90, int blocks_per_elem_f, int which, size_t block) {
  Ctx &c = C();
  std::vector<double> tmp((size_t)c.nelem * block);
  for (int ie = 0; ie < c.nelem; ++ie)
    memcpy(&tmp[(size_t)ie * block], f90 + ((size_t)ie * blocks_per_elem_f + which) * block, block * sizeof(double));
  h2d(dev, tmp.data(), tmp.size() * sizeof(double));
}
void push_levels(double *f90, const double *dev, int blocks_per_elem_f, int which, size_t block) {
  Ctx &c = C();
  std::vector<double> tmp((size_t)c.nelem * block);
  d2h(tmp.data(), dev, tmp.size() * sizeof(double));
  for (int ie = 0; ie < c.nelem; ++ie)
    memcpy(f90 + ((size_t)ie * blocks_per_elem_f + which) * block, &tmp[(size_t)ie * block], block * sizeof(double));
}

// tracer arrays: Fortran [ie][tl][qsize_d][lev][pt] vs device [ie][qsize][lev][pt]
void pull_tracers(double *dev, const double *f90, int ntl_f, int tl) {
  Ctx &c = C();
  const size_t e3 = (size_t)c.nlev * 16;
  std::vector<double> tmp((size_t)c.nelem * c.qsize * e3);
  for (int ie = 0; ie < c.nelem; ++ie)
    for (int q = 0; q < c.qsize; ++q)
      memcpy(&tmp[((size_t)ie * c.qsize + q) * e3], f90 + (((size_t)ie * ntl_f + tl) * c.qsize_d + q) * e3,
             e3 * sizeof(double));
  if (!tmp.empty()) h2d(dev, tmp.data(), tmp.size() * sizeof(double));
}
void push_tracers(double *f90, const double *dev, int ntl_f, int tl) {
  Ctx &c = C();
  const size_t e3 = (size_t)c.nlev * 16;
  std::vector<double> tmp((size_t)c.nelem * c.qsize * e3);
  if (!tmp.empty()) d2h(tmp.data(), dev, tmp.size() * sizeof(double));
  for (int ie = 0; ie < c.nelem; ++ie)
    for (int q = 0; q < c.qsize; ++q)
      memcpy(f90 + (((size_t)ie * ntl_f + tl) * c.qsize_d + q) * e3, &tmp[((size_t)ie * c.qsize + q) * e3],
             e3 * sizeof(double));
}

// Tracers::push_qdp (Tracers.cpp:46-51): both Qdp time levels of every tracer
// slot into the Fortran array (np,np,nlev,qsize_d,2,nelemd); the slots beyond
// qsize go back as they were received at init.
void push_qdp(double *Qdp) {
  Ctx &c = C();
  const size_t e3 = (size_t)c.nlev * 16;
  for (int tl = 0; tl < 2; ++tl) push_tracers(Qdp, c.qdp[tl], 2, tl);
  if (c.qsize_d > c.qsize && !c.qdp_tail.empty()) {
    const size_t tail = (size_t)(c.qsize_d - c.qsize) * e3;
    for (int ie = 0; ie < c.nelem; ++ie)
      for (int tl = 0; tl < 2; ++tl)
        memcpy(Qdp + (((size_t)ie * 2 + tl) * c.qsize_d + c.qsize) * e3, &c.qdp_tail[((size_t)ie * 2 + tl) * tail],
               tail * sizeof(double));
  }
}

// CAM forcing buffers FT, FM, FQ, allocated (zeroed, as the reference's
// views are) the first time a forcing run or a forcing push needs them
bool ensure_forcing() {
  Ctx &c = C();
  if (c.ft) return true;
  const size_t e3 = (size_t)c.nelem * c.nlev * 16;
  c.ft = dalloc<double>(e3);
  c.fm = dalloc<double>(2 * e3);
  c.fq = dalloc<double>(e3 * (c.qsize ? c.qsize : 1));
  return c.ft && c.fm && c.fq;
}

// ------------------------------------------------------------ diagnostics
// Diagnostics.cpp:37-185.  The device sums land in a staging buffer; slot
// ivar of each Fortran array is written with strided copies (the other slots
// keep their values, as in the reference).
double *diag_staging() {
  Ctx &c = C();
  if (!c.dg_dev) c.dg_dev = dalloc<double>((size_t)c.nelem * 16 * (4 + 2 * (size_t)c.qsize));
  return c.dg_dev;
}

// slot ivar of a Fortran array [ie][nslot][ncol] from a device array
// [ie][ncol] (ncol doubles per element, the first ncol of a row of width
// row_f doubles in the Fortran array)
void diag_copy(double *f90, int nslot, int ivar, size_t row_f, const double *dev, size_t ncol) {
  Ctx &c = C();
  hipMemcpy2DAsync(f90 + (size_t)ivar * row_f, (size_t)nslot * row_f * sizeof(double), dev, ncol * sizeof(double),
                   ncol * sizeof(double), c.nelem, hipMemcpyDeviceToHost, c.st);
}

DiagArgs diag_args(double *out) {
  Ctx &c = C();
  DiagArgs a{};
  a.nelem = c.nelem;
  a.nlev = c.nlev;
  a.qsize = c.qsize;
  a.use_cpstar = c.p.use_cpstar && c.qsize > 0;
  a.phis = c.phis;
  a.Q = c.Q;
  a.hyai = c.hyai_d;
  a.hybi = c.hybi_d;
  a.ps0 = c.ps0;
  a.out = out;
  return a;
}

// prim_diag_scalars (Diagnostics.cpp:37-88): Q to the Fortran Q array, then
// per column and tracer sum(Qdp*Q) -> qvar[slot], sum(Qdp) -> qmass[slot] and
// q1mass, at tracer level n0_qdp (before the advance) or np1_qdp (after)
void diag_scalars(bool before_advance, int ivar) {
  Ctx &c = C();
  if (!c.dg_set || c.p.time_step_type <= 0) return;
  double *out = diag_staging();
  const size_t e3 = (size_t)c.nlev * 16, nq = (size_t)c.nelem * c.qsize * 16;
  hipStreamSynchronize(c.st);
  push_tracers(c.dg[0], c.Q, 1, 0);
  for (int ie = 0; ie < c.nelem; ++ie)
    for (int q = c.qsize; q < c.qsize_d; ++q) memset(c.dg[0] + ((size_t)ie * c.qsize_d + q) * e3, 0, e3 * sizeof(double));
  if (c.qsize == 0) return;
  DiagArgs a = diag_args(out);
  a.qdp = c.qdp[before_advance ? c.n0_qdp : c.np1_qdp];
  launch_diag_scalars(c.st, a);
  const size_t row = (size_t)c.qsize_d * 16, ncol = (size_t)c.qsize * 16;
  diag_copy(c.dg[1], 4, ivar, row, out, ncol);       // qvar
  diag_copy(c.dg[2], 4, ivar, row, out + nq, ncol);  // qmass
  diag_copy(c.dg[3], 1, 0, row, out + nq, ncol);     // q1mass
  hipok(hipStreamSynchronize(c.st), "prim_diag_scalars");
}

// prim_energy_halftimes (Diagnostics.cpp:90-185) at dynamics level n0 (before
// the advance) or np1 (after): iener, kener, pener [slot] and iener_wet
void diag_energy(bool before_advance, int ivar) {
  Ctx &c = C();
  if (!c.dg_set) return;
  double *out = diag_staging();
  const size_t n = (size_t)c.nelem * 16;
  const int tl = before_advance ? c.n0 : c.np1;
  DiagArgs a = diag_args(out);
  a.v = c.v[tl];
  a.t = c.t[tl];
  a.ps = c.ps[tl];
  a.qdp = c.qdp[before_advance ? c.n0_qdp : c.np1_qdp];
  launch_diag_energy(c.st, a);
  diag_copy(c.dg[4], 4, ivar, 16, out, 16);          // iener
  diag_copy(c.dg[5], 1, 0, 16, out + n, 16);         // iener_wet
  diag_copy(c.dg[6], 4, ivar, 16, out + 2 * n, 16);  // kener
  diag_copy(c.dg[7], 4, ivar, 16, out + 3 * n, 16);  // pener
  hipok(hipStreamSynchronize(c.st), "prim_energy_halftimes");
}

// One dynamics step of the differentiable map from x into the stage states
// s, with the kernels of the forward step (no accumulators for the tracers)
void dyn_forward_stages(const StateBuf &x, Stages &s, double dt) {
  Ctx &c = C();
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  const int nsub = c.p.hypervis_subcycle;
  std::vector<StateBuf> &U = s.U, &H = s.H;
  copy_state(U[0], x, f3);
  caar_run(U[0], U[0], U[1], dt / 5.0, 0.0);
  caar_run(U[0], U[1], U[2], dt / 5.0, 0.0);
  caar_run(U[0], U[2], U[3], dt / 3.0, 0.0);
  caar_run(U[0], U[3], U[4], 2.0 * dt / 3.0, 0.0);
  copy_state(s.W, U[1], f3);
  launch_rk_combo(c.st, c.nelem, c.nlev, s.W.v, s.W.t, s.W.dp, U[0].v, U[0].t, U[0].dp);
  caar_run(s.W, U[4], H[0], 3.0 * dt / 4.0, 0.0);
  for (int cyc = 0; cyc < nsub; ++cyc) {
    copy_state(H[cyc + 1], H[cyc], f3);
    launch_hv_cycle(c.st, hv_args(H[cyc + 1], dt, 0.0));
  }
}

void adj_dyn_forward(AdjBufs &b, int i, double dt) { dyn_forward_stages(b.X[i], b.cur, dt); }

// X[0] from (v, T, ps), then the nd dynamics steps of one
// prim_run_subcycle_c call; H[nsub] holds the last step's output
void adj_forward_steps(AdjBufs &b, int nd, double dt, const double *v, const double *T, const double *ps) {
  Ctx &c = C();
  const size_t f3 = b.f3;
  hipMemcpyAsync(b.X[0].v, v, 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(b.X[0].t, T, f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  launch_dp3d_from_ps(c.st, c.nelem, c.nlev, c.hyai_d, c.hybi_d, c.ps0, ps, b.X[0].dp);
  for (int i = 0; i < nd; ++i) {
    adj_dyn_forward(b, i, dt);
    if (i + 1 < nd) copy_state(b.X[i + 1], b.cur.H[c.p.hypervis_subcycle], f3);
  }
}

// the configurations the differentiable form covers
bool adj_supported(const char *who) {
  Ctx &c = C();
  if (!ready(who)) return false;
  if (c.p.moist || !c.consthv || c.form != 0 || c.p.ftype != -1) {
    fail(70, "%s: needs a dry run, constant-coefficient hyperviscosity, no forcing and the FP64 form", who);
    return false;
  }
  if (!c.be_set) init_boundary_exchanges_c();
  return true;
}

// The reverse sweep of one prim_run_subcycle_c call: on entry lam_* hold
// the adjoint of the call's output (v, T, ps), on return that of its input.
// stages(i) returns the stage states of dynamics step i (recomputed or
// stored); the last step's are asked for first.
template <class F>
void adj_reverse_call(AdjBufs &b, int nd, double dt, F stages, double *lam_v, double *lam_T, double *lam_ps) {
  Ctx &c = C();
  const size_t f3 = b.f3;
  const int nsub = c.p.hypervis_subcycle;
  hipMemcpyAsync(b.LX.v, lam_v, 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(b.LX.t, lam_T, f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  for (int i = nd - 1; i >= 0; --i) {
    const Stages &st = stages(i);
    const std::vector<StateBuf> &U = st.U, &H = st.H;
    if (i == nd - 1) {
      if (c.p.rsplit == 0) {
        launch_adj_fill_dp(c.st, c.nelem, c.nlev, lam_ps, b.LX.dp);  // ps = hyai(0) ps0 + sum_k dp3d
      } else {
        // linearised PPM remap of v and T from the thicknesses of the last step
        AdjRemapArgs r{};
        r.nelem = c.nelem;
        r.nlev = c.nlev;
        r.alg = c.p.remap_alg;
        r.v = H[nsub].v;
        r.t = H[nsub].t;
        r.dp = H[nsub].dp;
        r.l_v = b.LX.v;
        r.l_t = b.LX.t;
        r.l_dp = b.LX.dp;
        r.l_ps = lam_ps;
        r.hyai = c.hyai_d;
        r.hybi = c.hybi_d;
        r.ps0 = c.ps0;
        r.scr = b.rscr;
        launch_adj_remap(c.st, r);
      }
    } else {
      copy_state(b.LX, b.L0, f3);  // adjoint of its output = adjoint of the next step's input
    }
    for (int cyc = nsub - 1; cyc >= 0; --cyc) hv_adjoint(b, H[cyc], H[cyc + 1], b.LX, dt / nsub);
    launch_adj_zero(c.st, 4 * f3, b.L0.v);
    launch_adj_zero(c.st, 4 * f3, b.L1.v);
    // stage 5 from w = (5 u1 - u0)/4 and u4; then stages 4..1
    caar_adjoint(b, U[4], b.LX, 3.0 * dt / 4.0, b.LX, false, b.L1, 1.25, &b.L0, -0.25);
    caar_adjoint(b, U[3], b.LX, 2.0 * dt / 3.0, b.LX, false, b.L0, 1.0, nullptr, 0.0);
    caar_adjoint(b, U[2], b.LX, dt / 3.0, b.LX, false, b.L0, 1.0, nullptr, 0.0);
    caar_adjoint(b, U[1], b.LX, dt / 5.0, b.L1, true, b.L0, 1.0, nullptr, 0.0);
    caar_adjoint(b, U[0], b.L1, dt / 5.0, b.L0, true, b.L0, 1.0, nullptr, 0.0);
  }
  hipMemcpyAsync(lam_v, b.L0.v, 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(lam_T, b.L0.t, f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  launch_adj_ps(c.st, c.nelem, c.nlev, c.hybi_d, b.L0.dp, lam_ps);  // dp3d = dhyai ps0 + dhybi ps
}

// The adjoint step on the buffers b (adj_layout for nd dynamics steps): the
// forward pass keeps each dynamics step's input and the last step's stage
// states; the reverse sweep recomputes the stage states of the others.
int adj_step_run(AdjBufs &b, int nd, double dt, const double *v, const double *T, const double *ps, double *lam_v,
                 double *lam_T, double *lam_ps, const char *who) {
  Ctx &c = C();
  {
    Scope s("adj_forward");
    adj_forward_steps(b, nd, dt, v, T, ps);
  }
  Scope s("adj_reverse");
  auto stages = [&](int i) -> const Stages & {
    if (i < nd - 1) adj_dyn_forward(b, i, dt);
    return b.cur;
  };
  adj_reverse_call(b, nd, dt, stages, lam_v, lam_T, lam_ps);
  return hipok(hipStreamSynchronize(c.st), who) ? 0 : 4;
}

// stage states for nsets dynamics steps (hommexx_hip_adjoint_segment)
bool seg_alloc(int nsets) {
  Ctx &c = C();
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  const int nsub = c.p.hypervis_subcycle;
  if (segb && segb->f3 == f3 && segb->nsub == nsub && segb->nsets >= nsets) return true;
  delete segb;
  segb = new SegBufs();
  SegBufs &g = *segb;
  const int per = 5 + 1 + nsub + 1;
  double *q = dalloc<double>(((size_t)nsets * per + 1) * 4 * f3 + (size_t)c.nelem * 16);
  if (!q) {
    delete segb;
    segb = nullptr;
    return false;
  }
  g.f3 = f3;
  g.nsub = nsub;
  g.nsets = nsets;
  g.sets.resize(nsets);
  for (Stages &st : g.sets) {
    for (int i = 0; i < 5; ++i) st.U.push_back(carve_state(q, f3));
    st.W = carve_state(q, f3);
    for (int i = 0; i <= nsub; ++i) st.H.push_back(carve_state(q, f3));
  }
  g.x = carve_state(q, f3);
  g.ps = q;
  return true;
}

}  // namespace

// =================================================================== C-ABI
extern "C" {

// Hommexx_Session.cpp:35
void initialize_hommexx_session() {
  Ctx &c = C();
  if (c.session) return;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    fail(1, "initialize_hommexx_session: no HIP device");
    return;
  }
  if (!hipok(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking), "hipStreamCreate")) return;
  if (!scratch_setup()) return;  // the no-allocation adjoint step's array on this device
  c.session = true;
}

// Hommexx_Session.cpp:75
void finalize_hommexx_session() {
  if (!g) return;
  free_all();
  if (g->st) hipStreamDestroy(g->st);
  const bool abort_flag = g->abort_on_error;
  delete g;
  g = nullptr;
  C().abort_on_error = abort_flag;
}

// mpi/mpi_cxx_f90_interface.cpp:20 (one rank: the communicator is recorded only)
void reset_cxx_comm(const int &f_comm) { (void)f_comm; }

// mpi/mpi_cxx_f90_interface.cpp:27
void init_connectivity(const int &num_local_elems) {
  Ctx &c = C();
  if (num_local_elems < 0) {
    fail(30, "init_connectivity: negative element count");
    return;
  }
  c.conn_nelem = num_local_elems;
  c.conn_final = false;
  c.conn_rl.assign((size_t)num_local_elems * 8, -1);
  c.conn_rp.assign((size_t)num_local_elems * 8, -1);
}

// mpi/mpi_cxx_f90_interface.cpp:34-55 (1-based ids, Fortran positions W,E,S,N,SW,SE,NW,NE)
void add_connection(const int &first_lid, const int &first_gid, const int &first_pos, const int &first_pid,
                    const int &second_lid, const int &second_gid, const int &second_pos, const int &second_pid) {
  Ctx &c = C();
  (void)first_gid;
  (void)second_gid;
  if (first_lid <= 0 || first_gid <= 0 || first_pos <= 0 || first_pid <= 0 || second_lid <= 0 || second_gid <= 0 ||
      second_pos <= 0 || second_pid <= 0) {
    fail(31, "add_connection: Fortran indices must start at 1");
    return;
  }
  if (first_pid != 1) return;  // only connections owned by this (single) rank are stored
  if (second_pid != 1) {
    fail(32, "add_connection: remote ranks are not supported on one GPU");
    return;
  }
  auto conv = [](int f) { return f <= 4 ? ((f - 1) + 2) % 4 : f - 1; };
  const int p1 = conv(first_pos), p2 = conv(second_pos);
  if (p1 > 7 || p2 > 7 || (p1 < 4) != (p2 < 4) || first_lid > c.conn_nelem || second_lid > c.conn_nelem) {
    fail(33, "add_connection: invalid connection (%d,%d)->(%d,%d)", first_lid, first_pos, second_lid, second_pos);
    return;
  }
  c.conn_rl[(size_t)(first_lid - 1) * 8 + p1] = second_lid - 1;
  c.conn_rp[(size_t)(first_lid - 1) * 8 + p1] = p2;
}

// mpi/mpi_cxx_f90_interface.cpp:57
void finalize_connectivity() {
  Ctx &c = C();
  for (int ie = 0; ie < c.conn_nelem; ++ie) {
    int missing = 0;
    for (int p = 0; p < 8; ++p) {
      if (c.conn_rl[(size_t)ie * 8 + p] < 0) {
        if (p < 4) {
          fail(34, "finalize_connectivity: element %d has no edge neighbour at position %d", ie + 1, p);
          return;
        }
        ++missing;
      }
    }
    if (missing > 1) {
      fail(35, "finalize_connectivity: element %d misses %d corner neighbours", ie + 1, missing);
      return;
    }
  }
  c.conn_final = true;
}

// cxx_f90_interface.cpp:207
void init_derivative_c(const double *const &dvv) {
  Ctx &c = C();
  memcpy(c.dvv, dvv, sizeof(c.dvv));
  c.dvv_set = true;
}

// cxx_f90_interface.cpp:33-117
void init_simulation_params_c(const int &remap_alg, const int &limiter_option, const int &rsplit, const int &qsplit,
                              const int &time_step_type, const int &energy_fixer, const int &qsize,
                              const int &state_frequency, const double &nu, const double &nu_p, const double &nu_q,
                              const double &nu_s, const double &nu_div, const double &nu_top,
                              const int &hypervis_order, const int &hypervis_subcycle, const double &hypervis_scaling,
                              const int &ftype, const bool &prescribed_wind, const bool &moisture,
                              const bool &disable_diagnostics, const bool &use_cpstar,
                              const bool &use_semi_lagrangian_transport) {
  Ctx &c = C();
  auto bad = [](const char *opt, const char *val) {
    fail(40, "init_simulation_params_c: unsupported option %s = %s", opt, val);
  };
  char buf[64];
  if (remap_alg != 1 && remap_alg != 2) return snprintf(buf, 64, "%d", remap_alg), bad("vert_remap_q_alg", buf);
  if (prescribed_wind) return bad("prescribed_wind", "true");
  if (hypervis_order != 2) return snprintf(buf, 64, "%d", hypervis_order), bad("hypervis_order", buf);
  if (use_semi_lagrangian_transport) return bad("use_semi_lagrangian_transport", "true");
  if (time_step_type != 5) return snprintf(buf, 64, "%d", time_step_type), bad("time_step_type", buf);
  if (limiter_option != 8 && limiter_option != 9)
    return snprintf(buf, 64, "%d", limiter_option), bad("limiter_option", buf);
  if (ftype != -1 && ftype != 0 && ftype != 2) return snprintf(buf, 64, "%d", ftype), bad("ftype", buf);
  if (!(nu_p > 0)) return snprintf(buf, 64, "%g", nu_p), bad("nu_p", buf);
  if (!(nu > 0)) return snprintf(buf, 64, "%g", nu), bad("nu", buf);
  if (!(nu_div > 0)) return snprintf(buf, 64, "%g", nu_div), bad("nu_div", buf);
  if (rsplit < 0) return snprintf(buf, 64, "%d", rsplit), bad("rsplit", buf);
  if (qsplit < 1 || hypervis_subcycle < 1 || qsize < 0)
    return bad("qsplit/hypervis_subcycle/qsize", "out of range");
  Params &p = c.p;
  p.remap_alg = remap_alg;
  p.limiter_option = limiter_option;
  p.rsplit = rsplit;
  p.qsplit = qsplit;
  p.time_step_type = time_step_type;
  p.energy_fixer = energy_fixer;
  p.qsize = qsize;
  p.state_frequency = state_frequency;
  p.nu = nu;
  p.nu_p = nu_p;
  p.nu_q = nu_q;
  p.nu_s = nu_s;
  p.nu_div = nu_div;
  p.nu_top = nu_top;
  p.hypervis_order = hypervis_order;
  p.hypervis_subcycle = hypervis_subcycle;
  p.hypervis_scaling = hypervis_scaling;
  p.ftype = ftype;
  p.prescribed_wind = prescribed_wind;
  p.moist = moisture;
  p.disable_diagnostics = disable_diagnostics;
  p.use_cpstar = use_cpstar;
  p.use_sl = use_semi_lagrangian_transport;
  if (p.nu != p.nu_div) {
    const double ratio = p.nu_div / p.nu;
    if (p.hypervis_scaling != 0.0) {
      p.nu_ratio1 = ratio * ratio;
      p.nu_ratio2 = 1.0;
    } else {
      p.nu_ratio1 = ratio;
      p.nu_ratio2 = ratio;
    }
  } else {
    p.nu_ratio1 = 1.0;
    p.nu_ratio2 = 1.0;
  }
  p.set = true;
}

// cxx_f90_interface.cpp:119 (hyam/hybm are accepted but unused, as in the reference)
void init_hvcoord_c(const double &ps0, const double *const &hybrid_am, const double *const &hybrid_ai,
                    const double *const &hybrid_bm, const double *const &hybrid_bi) {
  Ctx &c = C();
  (void)hybrid_am;
  (void)hybrid_bm;
  if (c.nlev <= 0) c.nlev = c.nlev_cfg;
  const int L = c.nlev;
  c.ps0 = ps0;
  c.hyai.assign(hybrid_ai, hybrid_ai + L + 1);
  c.hybi.assign(hybrid_bi, hybrid_bi + L + 1);
  c.hyai0 = hybrid_ai[0];
  std::vector<double> dp0(L);
  for (int k = 0; k < L; ++k) {
    const double da = c.hyai[k + 1] - c.hyai[k], db = c.hybi[k + 1] - c.hybi[k];
    dp0[k] = da * ps0 + db * ps0;
  }
  if (!c.hyai_d) {
    c.hyai_d = dalloc<double>(L + 1);
    c.hybi_d = dalloc<double>(L + 1);
    c.dp0_d = dalloc<double>(L);
  }
  h2d(c.hyai_d, c.hyai.data(), (L + 1) * sizeof(double));
  h2d(c.hybi_d, c.hybi.data(), (L + 1) * sizeof(double));
  h2d(c.dp0_d, dp0.data(), L * sizeof(double));
  c.hv_set = true;
}

// cxx_f90_interface.cpp:224-233
void init_elements_2d_c(const int &num_elems, const double *const &D, const double *const &Dinv,
                        const double *const &fcor, const double *const &mp, const double *const &spheremp,
                        const double *const &rspheremp, const double *const &metdet, const double *const &metinv,
                        const double *const &phis, const double *const &tensorvisc,
                        const double *const &vec_sph2cart, const bool &consthv) {
  Ctx &c = C();
  if (!c.session) initialize_hommexx_session();
  if (!c.p.set) {
    fail(50, "init_elements_2d_c: init_simulation_params_c must come first");
    return;
  }
  if (c.elems_set) {
    fail(51, "init_elements_2d_c: elements already initialised");
    return;
  }
  c.nelem = num_elems;
  c.nlev = c.nlev_cfg;
  c.qsize = c.p.qsize;
  c.qsize_d = c.qsize_d_cfg > 0 ? c.qsize_d_cfg : std::max(c.qsize, 1);
  if (c.qsize_d < c.qsize) {
    fail(52, "init_elements_2d_c: qsize %d exceeds qsize_d %d", c.qsize, c.qsize_d);
    return;
  }
  const size_t ne = num_elems;
  auto up = [&](double *&dst, const double *src, size_t n) {
    dst = dalloc<double>(n);
    if (src) h2d(dst, src, n * sizeof(double));
  };
  up(c.D, D, ne * 64);
  up(c.Dinv, Dinv, ne * 64);
  up(c.metinv, metinv, ne * 64);
  up(c.fcor, fcor, ne * 16);
  up(c.mp, mp, ne * 16);
  up(c.spheremp, spheremp, ne * 16);
  up(c.rspheremp, rspheremp, ne * 16);
  up(c.metdet, metdet, ne * 16);
  up(c.phis, phis, ne * 16);
  c.consthv = consthv;
  if (!consthv) {
    up(c.tensorvisc, tensorvisc, ne * 64);
    up(c.vec_sph2cart, vec_sph2cart, ne * 96);
  }
  c.dvv_d = dalloc<double>(16);
  h2d(c.dvv_d, c.dvv, sizeof(c.dvv));
  const size_t e3 = ne * c.nlev * 16;
  for (int tl = 0; tl < 3; ++tl) {
    c.v[tl] = dalloc<double>(2 * e3);
    c.t[tl] = dalloc<double>(e3);
    c.dp[tl] = dalloc<double>(e3);
    c.ps[tl] = dalloc<double>(ne * 16);
  }
  for (int tl = 0; tl < 2; ++tl) c.qdp[tl] = dalloc<double>(e3 * (c.qsize ? c.qsize : 1));
  c.Q = dalloc<double>(e3 * (c.qsize ? c.qsize : 1));
  c.phi = dalloc<double>(e3);
  c.omega = dalloc<double>(e3);
  c.eta_dot = dalloc<double>(e3);
  c.dpdiss_ave = dalloc<double>(e3);
  c.dpdiss_bih = dalloc<double>(e3);
  c.derived_dp = dalloc<double>(e3);
  c.divdp = dalloc<double>(e3);
  c.divdp_proj = dalloc<double>(e3);
  c.vn0 = dalloc<double>(2 * e3);
  c.ft = c.fm = c.fq = nullptr;  // CAM forcing buffers: on first use (ensure_forcing)
  c.elems_set = true;
}

// cxx_f90_interface.cpp:235-251
void init_elements_states_c(const double *const &v, const double *const &T, const double *const &dp3d,
                            const double *const &qdp, const double *const &ps_v) {
  Ctx &c = C();
  if (!c.elems_set) {
    fail(53, "init_elements_states_c: init_elements_2d_c must come first");
    return;
  }
  const size_t e3 = (size_t)c.nlev * 16;
  for (int tl = 0; tl < 3; ++tl) {
    pull_levels(c.v[tl], v, 3, tl, 2 * e3);
    pull_levels(c.t[tl], T, 3, tl, e3);
    pull_levels(c.dp[tl], dp3d, 3, tl, e3);
    pull_levels(c.ps[tl], ps_v, 3, tl, 16);
  }
  for (int tl = 0; tl < 2; ++tl) pull_tracers(c.qdp[tl], qdp, 2, tl);
  // tracers beyond qsize are carried unchanged (the reference keeps them in its qdp view)
  c.qdp_tail.clear();
  if (c.qsize_d > c.qsize) {
    const size_t tail = (size_t)(c.qsize_d - c.qsize) * e3;
    c.qdp_tail.resize((size_t)c.nelem * 2 * tail);
    for (int ie = 0; ie < c.nelem; ++ie)
      for (int tl = 0; tl < 2; ++tl)
        memcpy(&c.qdp_tail[((size_t)ie * 2 + tl) * tail], qdp + (((size_t)ie * 2 + tl) * c.qsize_d + c.qsize) * e3,
               tail * sizeof(double));
  }
  c.states_set = true;
}

// cxx_f90_interface.cpp:253-262, Diagnostics.cpp:17-35: the Fortran arrays
// prim_run_subcycle_c fills when diagnostics are enabled
void init_diagnostics_c(double *&q, double *&qvar, double *&qmass, double *&q1mass, double *&iener,
                        double *&iener_wet, double *&kener, double *&pener) {
  Ctx &c = C();
  double *p[8] = {q, qvar, qmass, q1mass, iener, iener_wet, kener, pener};
  bool all = true;
  for (int i = 0; i < 8; ++i) {
    c.dg[i] = p[i];
    all = all && p[i] != nullptr;
  }
  c.dg_set = all;
}

// cxx_f90_interface.cpp:264-280
void init_boundary_exchanges_c() {
  Ctx &c = C();
  if (!c.elems_set || !c.conn_final) {
    fail(54, "init_boundary_exchanges_c: elements and connectivity must be initialised first");
    return;
  }
  if (c.conn_nelem != c.nelem) {
    fail(21, "init_boundary_exchanges_c: connectivity has %d elements, element data %d", c.conn_nelem, c.nelem);
    return;
  }
  if (!c.dss_src) build_dss_tables();
  alloc_work();
}

// cxx_f90_interface.cpp:213-222 (1-based in)
void init_time_level_c(const int &nm1, const int &n0, const int &np1, const int &nstep, const int &nstep0) {
  Ctx &c = C();
  c.nm1 = nm1 - 1;
  c.n0 = n0 - 1;
  c.np1 = np1 - 1;
  c.nstep = nstep;
  c.nstep0 = nstep0;
}

// prim_driver.cpp:31-156: one remap interval.  Indices come back 0-based.
void prim_run_subcycle_c(const double &dt, int &nstep, int &nm1, int &n0, int &np1, const int &last_time_step) {
  Ctx &c = C();
  if (!ready("prim_run_subcycle_c")) return;
  if (!c.be_set) init_boundary_exchanges_c();
  const double dt_q = dt * c.p.qsplit;
  const double dt_remap = c.p.rsplit > 0 ? dt_q * c.p.rsplit : dt_q;
  // prim_driver.cpp:42-63: diagnostics at the state frequency, the first and
  // the last step, unless disabled (or no arrays were registered)
  const int nstep_end = c.nstep + c.p.qsplit * (c.p.rsplit > 0 ? c.p.rsplit : 1);
  const bool diag = !c.p.disable_diagnostics && c.dg_set &&
                    ((c.p.state_frequency > 0 && nstep_end % c.p.state_frequency == 0) || nstep_end == c.nstep0 ||
                     nstep_end >= last_time_step);
  update_tracers_levels();
  if (diag) {
    diag_scalars(true, 3);
    diag_energy(true, 2);
  }
  if ((c.p.ftype == 0 || c.p.ftype == 2) && !ensure_forcing()) return;
  if (c.p.ftype == 0) {
    Scope s("forcing");
    launch_state_forcing(c.st, c.nelem, c.nlev, dt_remap, c.ft, c.fm, c.t[c.n0], c.v[c.n0]);
    launch_tracer_forcing(c.st, c.nelem, c.nlev, c.qsize, c.p.moist, dt_remap, c.fq, c.qdp[c.n0_qdp], c.Q, c.hyai_d,
                          c.hybi_d, c.ps0, c.ps[c.n0]);
  } else if (c.p.ftype == 2) {
    Scope s("forcing");
    launch_state_forcing(c.st, c.nelem, c.nlev, dt_remap, c.ft, c.fm, c.t[c.n0], c.v[c.n0]);
  }
  if (diag) {
    diag_energy(true, 0);
    diag_scalars(true, 0);
  }
  {
    Scope s("dp3d_from_ps");
    launch_dp3d_from_ps(c.st, c.nelem, c.nlev, c.hyai_d, c.hybi_d, c.ps0, c.ps[c.n0], c.dp[c.n0]);
  }
  c.stopped = false;
  prim_step(dt);
  for (int r = 1; r < c.p.rsplit && !c.stopped; ++r) {
    leapfrog();
    prim_step(dt);
  }
  if (c.stopped) {
    hipStreamSynchronize(c.st);
    return;
  }
  update_tracers_levels();
  launch_zero_int(c.st, c.err_d);
  remap(dt_remap);
  if (diag) {
    diag_scalars(false, 1);
    diag_energy(false, 1);
  }
  leapfrog();
  nstep = c.nstep;
  nm1 = c.nm1;
  n0 = c.n0;
  np1 = c.np1;
  hipMemcpyAsync(c.err_h, c.err_d, sizeof(int), hipMemcpyDeviceToHost, c.st);
  if (!hipok(hipStreamSynchronize(c.st), "prim_run_subcycle_c")) return;
  if (*c.err_h & 1) fail(60, "Negative (or nan) layer thickness detected, aborting!");
  else if (*c.err_h & 2) fail(61, "vertical remap: Lagrangian levels deformed by more than one layer");
}

// cxx_f90_interface.cpp:126-154
void cxx_push_results_to_f90(double *&v, double *&T, double *&dp3d, double *&Qdp, double *&Q, double *&ps_v,
                             double *&omega_p) {
  Ctx &c = C();
  if (!c.states_set) {
    fail(55, "cxx_push_results_to_f90: no state");
    return;
  }
  hipStreamSynchronize(c.st);
  const size_t e3 = (size_t)c.nlev * 16;
  for (int tl = 0; tl < 3; ++tl) {
    push_levels(v, c.v[tl], 3, tl, 2 * e3);
    push_levels(T, c.t[tl], 3, tl, e3);
    push_levels(dp3d, c.dp[tl], 3, tl, e3);
    push_levels(ps_v, c.ps[tl], 3, tl, 16);
  }
  push_qdp(Qdp);
  push_tracers(Q, c.Q, 1, 0);
  for (int ie = 0; ie < c.nelem; ++ie)
    for (int q = c.qsize; q < c.qsize_d; ++q) memset(Q + ((size_t)ie * c.qsize_d + q) * e3, 0, e3 * sizeof(double));
  push_levels(omega_p, c.omega, 1, 0, e3);
}

// cxx_f90_interface.cpp:157-178
void cxx_push_forcing_to_f90(double *FM, double *FT, double *FQ) {
  Ctx &c = C();
  if (!ensure_forcing()) return;
  hipStreamSynchronize(c.st);
  const size_t e3 = (size_t)c.nlev * 16;
  push_levels(FM, c.fm, 1, 0, 2 * e3);
  push_levels(FT, c.ft, 1, 0, e3);
  if (c.p.ftype == 0 && FQ) {
    push_tracers(FQ, c.fq, 1, 0);
    // slots beyond qsize: what the last f90_push_forcing_to_cxx sent (zero
    // before any), as the reference's QSIZE_D-wide fq view holds them
    const size_t tail = (size_t)(c.qsize_d - c.qsize) * e3;
    for (int ie = 0; tail && ie < c.nelem; ++ie) {
      double *dst = FQ + ((size_t)ie * c.qsize_d + c.qsize) * e3;
      if (c.fq_tail.empty()) memset(dst, 0, tail * sizeof(double));
      else memcpy(dst, &c.fq_tail[(size_t)ie * tail], tail * sizeof(double));
    }
  }
}

// cxx_f90_interface.cpp:180-205.  FM, FT (and FQ with ftype 0) go to the
// device; Qdp goes the other way, as the reference's tracers.push_qdp does
// (:201): the Fortran array receives the device's Qdp.
void f90_push_forcing_to_cxx(double *FM, double *FT, double *FQ, double *Qdp) {
  Ctx &c = C();
  if (!ensure_forcing()) return;
  const size_t e3 = (size_t)c.nlev * 16;
  hipStreamSynchronize(c.st);
  pull_levels(c.fm, FM, 1, 0, 2 * e3);
  pull_levels(c.ft, FT, 1, 0, e3);
  if (c.p.ftype == 0 && FQ) {
    pull_tracers(c.fq, FQ, 1, 0);
    const size_t tail = (size_t)(c.qsize_d - c.qsize) * e3;
    c.fq_tail.resize((size_t)c.nelem * tail);
    for (int ie = 0; tail && ie < c.nelem; ++ie)
      memcpy(&c.fq_tail[(size_t)ie * tail], FQ + ((size_t)ie * c.qsize_d + c.qsize) * e3, tail * sizeof(double));
  }
  if (Qdp && c.states_set) push_qdp(Qdp);
}

// ============================================================ extensions
// Dimensions the reference fixes at compile time (PLEV, QSIZE_D); call before
// init_elements_2d_c.  Defaults: the build's PREQX_PLEV and QSIZE_D; a
// qsize_d of 0 or less means qsize.  nlev is at most REMAP_MAX_NLEV: the
// remap keeps each element's target table in LDS.
int hommexx_hip_configure(int nlev, int qsize_d) {
  Ctx &c = C();
  if (c.elems_set) return 1;
  if (nlev < 3 || nlev > REMAP_MAX_NLEV) return 2;
  c.nlev_cfg = nlev;
  c.qsize_d_cfg = qsize_d;
  return 0;
}

void hommexx_hip_dimensions(int *nlev, int *qsize_d) {
  Ctx &c = C();
  *nlev = c.nlev_cfg;
  *qsize_d = c.qsize_d_cfg;
}

void hommexx_hip_set_abort_on_error(int on) { C().abort_on_error = on != 0; }

int hommexx_hip_set_form(int form) {
  if (form != 0 && form != 1) return 1;
  C().form = form;
  return 0;
}
// Debugging aid: make prim_run_subcycle_c return after phase n (1: first RK
// stage, 2: all RK stages, 3: hyperviscosity, 4: first tracer stage,
// 5: tracers); 0 disables.  Time levels are not rotated when it stops.
void hommexx_hip_debug_stop(int phase) { C().debug_stop = phase; }
int hommexx_hip_last_error() { return C().last_error; }
const char *hommexx_hip_error_message() { return C().msg.c_str(); }
void hommexx_hip_clear_error() {
  C().last_error = 0;
  C().msg.clear();
}
void hommexx_hip_synchronize() {
  if (C().st) hipStreamSynchronize(C().st);
}
void *hommexx_hip_stream() { return (void *)C().st; }

// Per-phase device timing with HIP events on the library's stream.
void hommexx_hip_profile(int enable) {
  Ctx &c = C();
  if (c.prof && !enable) flush_timers();
  c.prof = enable != 0;
}
void hommexx_hip_profile_only(const char *name) {
  Ctx &c = C();
  c.prof_only = name ? name : "";
}
void hommexx_hip_profile_reset() {
  Ctx &c = C();
  flush_timers();
  c.timers.clear();
}
int hommexx_hip_profile_get(const char *name, int *count, double *total_ms) {
  Ctx &c = C();
  flush_timers();
  auto it = c.timers.find(name);
  if (it == c.timers.end()) {
    *count = 0;
    *total_ms = 0;
    return 1;
  }
  *count = it->second.count;
  *total_ms = it->second.total_ms;
  return 0;
}
// names of the timed phases, '\n'-separated, into buf (returns needed length)
int hommexx_hip_profile_names(char *buf, int len) {
  Ctx &c = C();
  flush_timers();
  std::string s;
  for (auto &kv : c.timers) {
    s += kv.first;
    s += "\n";
  }
  if (buf && len > 0) {
    strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)s.size() + 1;
}

// ---------------------------------------------------------- differentiable form
// Device copies of the prognostic state at time level n0, in the library's
// layouts: v [ie][lev][2][pt], T [ie][lev][pt], ps [ie][pt].  Used by the
// checkpointing harness.
int hommexx_hip_get_state(double *v, double *T, double *ps) {
  Ctx &c = C();
  if (!c.states_set) {
    fail(55, "hommexx_hip_get_state: no state");
    return 1;
  }
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  hipMemcpyAsync(v, c.v[c.n0], 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(T, c.t[c.n0], f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(ps, c.ps[c.n0], (size_t)c.nelem * 16 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  return hipok(hipStreamSynchronize(c.st), "hommexx_hip_get_state") ? 0 : 2;
}

// One field of the current state into device memory dst, library layout:
// which 0 = v [ie][lev][2][pt], 1 = T, 2 = dp3d [ie][lev][pt], 3 = ps_v
// [ie][pt] (time level n0), 4 = Qdp of tracer q [ie][lev][pt] (the tracer
// level the next step starts from, (nstep / qsplit) % 2), 5 = Q of tracer q.
int hommexx_hip_get_field(int which, int q, double *dst) {
  Ctx &c = C();
  if (!c.states_set) {
    fail(55, "hommexx_hip_get_field: no state");
    return 1;
  }
  if (which < 0 || which > 5 || (which >= 4 && (q < 0 || q >= c.qsize))) {
    fail(56, "hommexx_hip_get_field: no field %d (tracer %d)", which, q);
    return 1;
  }
  const size_t f3 = (size_t)c.nelem * c.nlev * 16, e3 = (size_t)c.nlev * 16;
  if (which <= 3) {
    const double *srcs[4] = {c.v[c.n0], c.t[c.n0], c.dp[c.n0], c.ps[c.n0]};
    const size_t n[4] = {2 * f3, f3, f3, (size_t)c.nelem * 16};
    hipMemcpyAsync(dst, srcs[which], n[which] * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  } else {
    const double *base = which == 4 ? c.qdp[(c.nstep / c.p.qsplit) % 2] : c.Q;
    hipMemcpy2DAsync(dst, e3 * sizeof(double), base + (size_t)q * e3, (size_t)c.qsize * e3 * sizeof(double),
                     e3 * sizeof(double), c.nelem, hipMemcpyDeviceToDevice, c.st);
  }
  return hipok(hipStreamSynchronize(c.st), "hommexx_hip_get_field") ? 0 : 2;
}

int hommexx_hip_set_state(const double *v, const double *T, const double *ps) {
  Ctx &c = C();
  if (!c.states_set || !c.hv_set) {
    fail(55, "hommexx_hip_set_state: no state");
    return 1;
  }
  const size_t f3 = (size_t)c.nelem * c.nlev * 16;
  hipMemcpyAsync(c.v[c.n0], v, 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(c.t[c.n0], T, f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(c.ps[c.n0], ps, (size_t)c.nelem * 16 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  launch_dp3d_from_ps(c.st, c.nelem, c.nlev, c.hyai_d, c.hybi_d, c.ps0, c.ps[c.n0], c.dp[c.n0]);
  return hipok(hipStreamSynchronize(c.st), "hommexx_hip_set_state") ? 0 : 2;
}

// Adjoint of one step of length dt (one prim_run_subcycle_c call) of a dry
// configuration (any rsplit and qsplit; constant-coefficient hyperviscosity;
// the FP64 form), taken at the step's input state (v, T, ps):
// on entry lam_* hold the adjoint of the step's output (v, T, ps), on return
// the adjoint of its input.  All pointers are device memory in the layouts of
// hommexx_hip_get_state.  The model state of the session is not changed.
int hommexx_hip_adjoint_step(double dt, const double *v, const double *T, const double *ps, double *lam_v,
                             double *lam_T, double *lam_ps) {
  Ctx &c = C();
  if (!adj_supported("hommexx_hip_adjoint_step")) return 2;
  // prim_run_subcycle_c: qsplit dynamics steps per prim_step, rsplit (at
  // least one) prim_steps, then the remap (prim_driver.cpp:115-130,
  // prim_step.cpp:57-80)
  const int nd = c.p.qsplit * (c.p.rsplit > 0 ? c.p.rsplit : 1);
  if (!adj_alloc(nd)) return 3;
  return adj_step_run(*adjb, nd, dt, v, T, ps, lam_v, lam_T, lam_ps, "hommexx_hip_adjoint_step");
}

// As hommexx_hip_adjoint_step, but every buffer of the step comes from the
// library's fixed device array `scratch` (ADJ_SCRATCH_DOUBLES doubles, one
// per GPU, held until the process ends) instead of device allocations: for
// callers that cannot allocate (graph capture, serving loops).  Larger
// configurations are rejected (error 71).  All calls on one GPU share its
// array, so calls that overlap on one GPU race.
int hommexx_hip_adjoint_step_noalloc(double dt, const double *v, const double *T, const double *ps, double *lam_v,
                                     double *lam_T, double *lam_ps) {
  Ctx &c = C();
  if (!ready("hommexx_hip_adjoint_step_noalloc")) return 2;
  if (!c.be_set) {
    fail(72, "hommexx_hip_adjoint_step_noalloc: init_boundary_exchanges_c has not run (it allocates)");
    return 2;
  }
  if (!adj_supported("hommexx_hip_adjoint_step_noalloc")) return 2;
  const int nd = c.p.qsplit * (c.p.rsplit > 0 ? c.p.rsplit : 1);
  const size_t need = adj_total_doubles(nd);
  if (need > ADJ_SCRATCH_DOUBLES) {
    fail(71, "hommexx_hip_adjoint_step_noalloc: needs %zu doubles of scratch, the library holds %zu", need,
         (size_t)ADJ_SCRATCH_DOUBLES);
    return 3;
  }
  double *base = scratch_pointer();
  if (!base) return 3;
  static AdjBufs nb;  // the layout is rebuilt only when the configuration changes
  if (nb.base != base || nb.f3 != (size_t)c.nelem * c.nlev * 16 || nb.nsub != c.p.hypervis_subcycle || nb.nd != nd)
    adj_layout(nb, base, nd);
  // the array keeps the last call's contents: start from zeros, as the
  // allocated buffers of hommexx_hip_adjoint_step do
  launch_adj_zero(c.st, need, base);
  return adj_step_run(nb, nd, dt, v, T, ps, lam_v, lam_T, lam_ps, "hommexx_hip_adjoint_step_noalloc");
}


// The adjoint of n consecutive prim_run_subcycle_c calls of length dt (a
// checkpointed segment), from the state (v, T, ps) before the first: the
// forward pass keeps the stage states of every dynamics step (40 F3 each at
// hypervis_subcycle 3), so the reverse sweep recomputes nothing.  lam_* as
// for hommexx_hip_adjoint_step (output adjoint in, input adjoint out).
int hommexx_hip_adjoint_segment(double dt, int n, const double *v, const double *T, const double *ps, double *lam_v,
                                double *lam_T, double *lam_ps) {
  Ctx &c = C();
  if (!adj_supported("hommexx_hip_adjoint_segment")) return 2;
  if (n < 1) {
    fail(73, "hommexx_hip_adjoint_segment: %d calls", n);
    return 2;
  }
  const int nd = c.p.qsplit * (c.p.rsplit > 0 ? c.p.rsplit : 1);
  if (!adj_alloc(nd) || !seg_alloc(n * nd)) return 3;
  AdjBufs &b = *adjb;
  SegBufs &g = *segb;
  const size_t f3 = b.f3;
  const int nsub = c.p.hypervis_subcycle;
  {
    Scope s("adj_forward");
    launch_zero_int(c.st, c.err_d);
    hipMemcpyAsync(g.x.v, v, 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
    hipMemcpyAsync(g.x.t, T, f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
    launch_dp3d_from_ps(c.st, c.nelem, c.nlev, c.hyai_d, c.hybi_d, c.ps0, ps, g.x.dp);
    for (int call = 0; call < n; ++call) {
      for (int i = 0; i < nd; ++i)
        dyn_forward_stages(i == 0 ? g.x : g.sets[call * nd + i - 1].H[nsub], g.sets[call * nd + i], dt);
      if (call + 1 == n) break;
      // the next call's input: remapped v, T (rsplit > 0), ps from the
      // thicknesses and dp3d from ps, as prim_run_subcycle_c leaves them
      copy_state(g.x, g.sets[call * nd + nd - 1].H[nsub], f3);
      if (c.p.rsplit > 0) {
        RemapArgs a{};
        a.nelem = c.nelem;
        a.nlev = c.nlev;
        a.qsize = 0;
        a.nstates = 3;
        a.remap_alg = c.p.remap_alg;
        a.v = g.x.v;
        a.t = g.x.t;
        a.dp = g.x.dp;
        a.ps = g.ps;
        a.hyai = c.hyai_d;
        a.hybi = c.hybi_d;
        a.ps0 = c.ps0;
        a.hyai0 = c.hyai0;
        a.error = c.err_d;
        launch_remap(c.st, a);
      } else {
        launch_ps_from_dp(c.st, c.nelem, c.nlev, c.hyai0, c.ps0, g.x.dp, g.ps);
      }
      launch_dp3d_from_ps(c.st, c.nelem, c.nlev, c.hyai_d, c.hybi_d, c.ps0, g.ps, g.x.dp);
    }
  }
  {
    Scope s("adj_reverse");
    for (int call = n - 1; call >= 0; --call) {
      auto stages = [&](int i) -> const Stages & { return g.sets[call * nd + i]; };
      adj_reverse_call(b, nd, dt, stages, lam_v, lam_T, lam_ps);
    }
  }
  hipMemcpyAsync(c.err_h, c.err_d, sizeof(int), hipMemcpyDeviceToHost, c.st);
  if (!hipok(hipStreamSynchronize(c.st), "hommexx_hip_adjoint_segment")) return 4;
  if (*c.err_h & 1) {
    fail(60, "Negative (or nan) layer thickness detected, aborting!");
    return 5;
  }
  return 0;
}

// The map hommexx_hip_adjoint_step differentiates, forward: (v, T, ps) is
// advanced in place by the dynamics steps of one prim_run_subcycle_c call and
// the remap of v and T (rsplit > 0), with no tracer step; for a dry run the
// result equals the (v, T, ps) that prim_run_subcycle_c produces.  The
// session's model state is not changed.
int hommexx_hip_dynamics_step(double dt, double *v, double *T, double *ps) {
  Ctx &c = C();
  if (!adj_supported("hommexx_hip_dynamics_step")) return 2;
  const int nd = c.p.qsplit * (c.p.rsplit > 0 ? c.p.rsplit : 1);
  if (!adj_alloc(nd)) return 3;
  AdjBufs &b = *adjb;
  const size_t f3 = b.f3;
  launch_zero_int(c.st, c.err_d);
  adj_forward_steps(b, nd, dt, v, T, ps);
  const StateBuf &o = b.cur.H[c.p.hypervis_subcycle];
  if (c.p.rsplit > 0) {
    RemapArgs a{};
    a.nelem = c.nelem;
    a.nlev = c.nlev;
    a.qsize = 0;
    a.nstates = 3;
    a.remap_alg = c.p.remap_alg;
    a.v = o.v;
    a.t = o.t;
    a.dp = o.dp;
    a.ps = ps;
    a.hyai = c.hyai_d;
    a.hybi = c.hybi_d;
    a.ps0 = c.ps0;
    a.hyai0 = c.hyai0;
    a.error = c.err_d;
    launch_remap(c.st, a);
  } else {
    launch_ps_from_dp(c.st, c.nelem, c.nlev, c.hyai0, c.ps0, o.dp, ps);
  }
  hipMemcpyAsync(v, o.v, 2 * f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(T, o.t, f3 * sizeof(double), hipMemcpyDeviceToDevice, c.st);
  hipMemcpyAsync(c.err_h, c.err_d, sizeof(int), hipMemcpyDeviceToHost, c.st);
  if (!hipok(hipStreamSynchronize(c.st), "hommexx_hip_dynamics_step")) return 4;
  if (*c.err_h & 1) {
    fail(60, "Negative (or nan) layer thickness detected, aborting!");
    return 5;
  }
  return 0;
}

// Current time-level indices (0-based) and sizes.
void hommexx_hip_time_levels(int *nm1, int *n0, int *np1, int *nstep, int *n0_qdp, int *np1_qdp) {
  Ctx &c = C();
  *nm1 = c.nm1;
  *n0 = c.n0;
  *np1 = c.np1;
  *nstep = c.nstep;
  *n0_qdp = c.n0_qdp;
  *np1_qdp = c.np1_qdp;
}

}  // extern "C"
