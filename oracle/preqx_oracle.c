/*
 * preqx_oracle.c -- CPU restatement of the HOMMEXX preqx time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is used by
 * tests/, by __graft_entry__.smoke() as the checker and by bench.py's CPU
 * baseline leg.  The product (libhommexx_amd.so) never links or calls it.
 *
 * It restates, in plain C with OpenMP over elements, the algorithm of the
 * reference C++ back end (paths relative to the reference root):
 *   step orchestration  src/share/cxx/prim_driver.cpp:31-206,
 *                       prim_step.cpp:20-104, prim_advance_exp.cpp:25-161,
 *                       prim_advec_tracers_remap.cpp:32-90, TimeLevel.hpp:37-67
 *   CAAR                CaarFunctorImpl.hpp:98-610 and the serial scans :621-650,
 *                       :689-729, :854-889
 *   sphere operators    SphereOperators.hpp:291-862
 *   hyperviscosity      HyperviscosityFunctorImpl.cpp:56-112,
 *                       HyperviscosityFunctorImpl.hpp:71-257
 *   Euler tracer step   EulerStepFunctorImpl.hpp:196-561 (team limiters :693-884)
 *   vertical remap      RemapFunctor.hpp:226-467, PpmRemap.hpp:78-673
 *   DSS                 mpi/BoundaryExchange.cpp:285-849, ConnectivityHelpers.hpp
 *   diagnostics         Diagnostics.cpp:37-185, called as prim_driver.cpp:49-63,
 *                       84-88, 140-144 place them
 * Floating-point operations are written in the reference's order (left to
 * right, serial reductions and scans in index order, as the reference does on
 * CPU and with HOMMEXX_GPU_BFB_WITH_CPU on GPU).  Build with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * How each part is pinned (tests/test_oracle.py, tests/test_oracle_pinning.py):
 *   sphere operators  the reference's golden vectors (bit-exact), closed forms
 *                     on the sphere, sum(weak laplacian) == 0
 *   DSS               continuity of every node, DSS(spheremp c) rspheremp == c,
 *                     54 ne^2 + 2 unique nodes
 *   CAAR              the steady JW jet stays balanced, converging with ne;
 *                     over uniform p and T one stage gives omega p = the
 *                     pressure tendency of its own thickness change and
 *                     d ln T = kappa d ln p (potential temperature kept)
 *   RK5 stages        third-order convergence in time (error / 8 per halving)
 *   hyperviscosity    damping of a Y2 harmonic at dt nu (6/a^2)^2 per subcycle;
 *                     nu_top sponge: Y2 damped at {4,2,1} dt nu_top 6/a^2 on
 *                     the top three levels, not below;
 *                     tensor form: identity tensor == simple laplacian bit for
 *                     bit, cartesian vector laplacian leaves rigid rotation
 *                     undamped and agrees with the contravariant form, both to
 *                     a truncation error converging at second order
 *   Euler step        constant mixing ratio kept; zonal tracer kept by the
 *                     zonal jet (third-order convergence); limiter bounds,
 *                     mass and 1-norm properties (preqx_ut.cpp:1335-1532)
 *   PPM remap         identity, constants, column mass, monotone range,
 *                     linear profiles exact in the interior
 *   CAM forcing       forced tracer and (moist) surface-pressure mass equal the
 *                     initial mass plus the clipped dt*FQ; a uniform FT raises
 *                     T by dt*FT
 *   full step         dry and tracer mass conserved to 1e-14
 *   diagnostics       PEner = phis (ps_v - hyai(0) ps0 + hyai(L) ps0) to
 *                     rounding; Qmass summed over the sphere conserved from
 *                     slot 3 (before) to slot 1 (after); Qvar = sum Qdp^2/dp;
 *                     IEner_wet = 0 without cp*; the slot schedule of
 *                     state_frequency, nstep0 and the last step
 * Each check is shown to fail on a build with one planted mistake
 * (ORACLE_PLANT below).  The exact accumulation order of the DSS follows
 * ConnectivityHelpers.hpp; property checks cannot see summation order, so
 * that order is pinned by reading only.
 *
 * Internal layout mirrors the reference C++ views: levels fastest,
 *   3-D scalar  [ie][igp][jgp][lev]
 *   2-D tensor  [ie][idim][jdim][igp][jgp] (flat copy of the Fortran arrays).
 * Entry points take and return the flat Fortran arrays of the C-ABI.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted mistakes for the pinning tests (tests/test_oracle_pinning.py):
 * built with -DORACLE_PLANT=n the oracle carries one deliberate error, which
 * the invariant tests must detect.  The default build (0) has none. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#define PLANT(n) (ORACLE_PLANT == (n))

#define NP 4
#define NP2 16
#define NCONN 8

static const double RGAS = 287.04;
static const double CP = 1005.0;
static const double RWATER_VAPOR = 461.5;
static const double CPWATER_VAPOR = 1870.0;
#define KAPPA (RGAS / CP)
#define RREARTH (1.0 / 6.376e6)

typedef struct {
  int nelem, nlev, qsize, qsize_d;
  /* simulation params */
  int remap_alg, limiter_option, rsplit, qsplit, time_step_type, state_frequency;
  int hypervis_order, hypervis_subcycle, ftype, moist, disable_diagnostics;
  double nu, nu_p, nu_q, nu_s, nu_div, nu_top, hypervis_scaling, nu_ratio1, nu_ratio2;
  int consthv;
  /* derivative: dvv[a][b] */
  double dvv[NP][NP];
  /* 2-D element data */
  double *D, *Dinv, *metinv, *tensorvisc, *vec_sph2cart;
  double *fcor, *mp, *spheremp, *rspheremp, *metdet, *phis;
  /* vertical coordinate */
  double ps0, hyai0;
  double *hyai, *hybi, *hyai_delta, *hybi_delta, *dp0;
  /* connectivity, reference C++ numbering (S,N,W,E,SW,SE,NW,NE) */
  int *conn_lid, *conn_pos, *conn_dir;
  /* prognostic state */
  double *v, *t, *dp3d, *ps_v, *qdp, *Q;
  /* derived */
  double *phi, *omega_p, *eta_dot_dpdn, *dpdiss_ave, *dpdiss_bih, *derived_dp, *divdp, *divdp_proj, *vn0;
  /* buffers */
  double *qtens_bih, *qlim, *vstar, *dpdissk, *dpbuf, *ttens, *dptens, *vtens;
  double *fm, *ft, *fq; /* CAM forcing, [ie][c][pt][lev], [ie][pt][lev], [ie][q][pt][lev] */
  double *dss_buf;   /* [nelem][8][NP][max fields * nlev] */
  int dss_buf_fields;
  /* time levels */
  int nm1, n0, np1, nstep, nstep0, n0_qdp, np1_qdp;
  /* diagnostics (Fortran elem%state%Q and elem%accum arrays, flat) */
  int diag_on, use_cpstar, last_time_step;
  double *dg_q, *dg_qvar, *dg_qmass, *dg_q1mass, *dg_iener, *dg_iener_wet, *dg_kener, *dg_pener;
  int error;
  int debug_stop, stopped;
} Oracle;

/* ------------------------------------------------------------------ layout */
#define L (o->nlev)
#define E3 ((size_t)NP2 * L)                 /* one element, one 3-D scalar */
#define IDX3(ie, ig, jg, k) ((size_t)(ie) * E3 + ((size_t)((ig) * NP + (jg))) * L + (k))
#define V(ie, tl, c) (o->v + (((size_t)(ie) * 3 + (tl)) * 2 + (c)) * E3)
#define T(ie, tl) (o->t + ((size_t)(ie) * 3 + (tl)) * E3)
#define DP(ie, tl) (o->dp3d + ((size_t)(ie) * 3 + (tl)) * E3)
#define PSV(ie, tl) (o->ps_v + ((size_t)(ie) * 3 + (tl)) * NP2)
#define QDP(ie, tl, q) (o->qdp + (((size_t)(ie) * 2 + (tl)) * o->qsize + (q)) * E3)
#define E2(a, ie) ((a) + (size_t)(ie) * NP2)
#define T2(a, ie) ((a) + (size_t)(ie) * 4 * NP2)  /* [2][2][NP][NP] */
#define AT(p, ig, jg, k) ((p)[((ig) * NP + (jg)) * L + (k)])
#define M2(p, ig, jg) ((p)[(ig) * NP + (jg)])
#define M4(p, a, b, ig, jg) ((p)[(((a) * 2 + (b)) * NP + (ig)) * NP + (jg)])

static double *alloc0(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ======================================================= sphere operators
 * Element-local, multi-level.  ``s``/``out`` point at one element's
 * [NP][NP][L] block; vector fields are two such blocks (stride E3).
 * ``nl`` is the number of levels processed (all, or the top 3 for nu_top). */
typedef struct { const Oracle *o; int ie; } Ctx;

static void gradient_sphere(const Oracle *o, int ie, int nl, const double *s, double *g0, double *g1) {
  const double *Dinv = T2(o->Dinv, ie);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        double v0 = 0, v1 = 0;
        for (int kk = 0; kk < NP; ++kk) {
          v0 += o->dvv[jg][kk] * AT(s, ig, kk, k);
          v1 += o->dvv[ig][kk] * AT(s, kk, jg, k);
        }
        v0 *= RREARTH;
        v1 *= RREARTH;
        AT(g0, ig, jg, k) = M4(Dinv, 0, 0, ig, jg) * v0 + M4(Dinv, 0, 1, ig, jg) * v1;
        AT(g1, ig, jg, k) = M4(Dinv, 1, 0, ig, jg) * v0 + M4(Dinv, 1, 1, ig, jg) * v1;
      }
}

static void gradient_sphere_update(const Oracle *o, int ie, const double *s, double *g0, double *g1) {
  const double *Dinv = T2(o->Dinv, ie);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < L; ++k) {
        double dsdx = 0, dsdy = 0;
        for (int kk = 0; kk < NP; ++kk) {
          dsdx += o->dvv[jg][kk] * AT(s, ig, kk, k);
          dsdy += o->dvv[ig][kk] * AT(s, kk, jg, k);
        }
        dsdx *= RREARTH;
        dsdy *= RREARTH;
        AT(g0, ig, jg, k) += M4(Dinv, 0, 0, ig, jg) * dsdx + M4(Dinv, 0, 1, ig, jg) * dsdy;
        AT(g1, ig, jg, k) += M4(Dinv, 1, 0, ig, jg) * dsdx + M4(Dinv, 1, 1, ig, jg) * dsdy;
      }
}

static void divergence_sphere(const Oracle *o, int ie, int nl, const double *v0, const double *v1,
                              double *div, double *gv0, double *gv1) {
  const double *Dinv = T2(o->Dinv, ie), *metdet = E2(o->metdet, ie);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double a = AT(v0, ig, jg, k), b = AT(v1, ig, jg, k);
        AT(gv0, ig, jg, k) = (M4(Dinv, 0, 0, ig, jg) * a + M4(Dinv, 1, 0, ig, jg) * b) * M2(metdet, ig, jg);
        AT(gv1, ig, jg, k) = (M4(Dinv, 0, 1, ig, jg) * a + M4(Dinv, 1, 1, ig, jg) * b) * M2(metdet, ig, jg);
      }
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        double dudx = 0, dvdy = 0;
        for (int kk = 0; kk < NP; ++kk) {
          dudx += (PLANT(8) ? o->dvv[kk][jg] : o->dvv[jg][kk]) * AT(gv0, ig, kk, k); /* plant 8: transposed dvv */
          dvdy += o->dvv[ig][kk] * AT(gv1, kk, jg, k);
        }
        AT(div, ig, jg, k) = (dudx + dvdy) * (1.0 / M2(metdet, ig, jg) * RREARTH);
      }
}

/* qtens = qdp + alpha*div(vstar*qdp) (+ qtens on input if add_hv) */
static void divergence_sphere_update(const Oracle *o, int ie, double alpha, int add_hv,
                                     const double *vs0, const double *vs1, const double *qdp,
                                     double *qtens, double *gv0, double *gv1) {
  const double *Dinv = T2(o->Dinv, ie), *metdet = E2(o->metdet, ie);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < L; ++k) {
        const double q = AT(qdp, ig, jg, k);
        const double a = AT(vs0, ig, jg, k) * q, b = AT(vs1, ig, jg, k) * q;
        AT(gv0, ig, jg, k) = (M4(Dinv, 0, 0, ig, jg) * a + M4(Dinv, 1, 0, ig, jg) * b) * M2(metdet, ig, jg);
        AT(gv1, ig, jg, k) = (M4(Dinv, 0, 1, ig, jg) * a + M4(Dinv, 1, 1, ig, jg) * b) * M2(metdet, ig, jg);
      }
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < L; ++k) {
        double dudx = 0, dvdy = 0;
        for (int kk = 0; kk < NP; ++kk) {
          dudx += (PLANT(8) ? o->dvv[kk][jg] : o->dvv[jg][kk]) * AT(gv0, ig, kk, k); /* plant 8: transposed dvv */
          dvdy += o->dvv[ig][kk] * AT(gv1, kk, jg, k);
        }
        const double q0 = add_hv ? AT(qtens, ig, jg, k) : 0;
        AT(qtens, ig, jg, k) = (AT(qdp, ig, jg, k) +
                                (PLANT(5) ? -alpha : alpha) * ((dudx + dvdy) * (1.0 / M2(metdet, ig, jg) * RREARTH)) +
                                q0); /* plant 5: flux-divergence sign */
      }
}

static void vorticity_sphere(const Oracle *o, int ie, int nl, const double *u, const double *v,
                             double *vort, double *c0, double *c1) {
  const double *D = T2(o->D, ie), *metdet = E2(o->metdet, ie);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double a = AT(u, ig, jg, k), b = AT(v, ig, jg, k);
        AT(c0, ig, jg, k) = M4(D, 0, 0, ig, jg) * a + M4(D, 0, 1, ig, jg) * b;
        AT(c1, ig, jg, k) = M4(D, 1, 0, ig, jg) * a + M4(D, 1, 1, ig, jg) * b;
      }
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        double dudy = 0, dvdx = 0;
        for (int kk = 0; kk < NP; ++kk) {
          dvdx += o->dvv[jg][kk] * AT(c1, ig, kk, k);
          dudy += o->dvv[ig][kk] * AT(c0, kk, jg, k);
        }
        AT(vort, ig, jg, k) = (dvdx - dudy) * (1.0 / M2(metdet, ig, jg) * RREARTH);
      }
}

/* weak divergence; v0/v1 are overwritten (contravariant, as in the reference) */
static void divergence_sphere_wk(const Oracle *o, int ie, int nl, double *v0, double *v1, double *div) {
  const double *Dinv = T2(o->Dinv, ie), *sph = E2(o->spheremp, ie);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double a = AT(v0, ig, jg, k), b = AT(v1, ig, jg, k);
        AT(v0, ig, jg, k) = M4(Dinv, 0, 0, ig, jg) * a + M4(Dinv, 1, 0, ig, jg) * b;
        AT(v1, ig, jg, k) = M4(Dinv, 0, 1, ig, jg) * a + M4(Dinv, 1, 1, ig, jg) * b;
      }
  for (int n = 0; n < NP; ++n)
    for (int m = 0; m < NP; ++m)
      for (int k = 0; k < nl; ++k) {
        double dd = 0;
        for (int j = 0; j < NP; ++j)
          dd -= (M2(sph, n, j) * AT(v0, n, j, k) * o->dvv[j][m] +
                 M2(sph, j, m) * AT(v1, j, m, k) * o->dvv[j][n]) * RREARTH;
        AT(div, n, m, k) = dd;
      }
}

static void laplace_simple(const Oracle *o, int ie, int nl, const double *f, double *lap, double *b0, double *b1) {
  gradient_sphere(o, ie, nl, f, b0, b1);
  divergence_sphere_wk(o, ie, nl, b0, b1, lap);
}

static void laplace_tensor(const Oracle *o, int ie, int nl, const double *f, double *lap,
                           double *b0, double *b1, double *s0, double *s1) {
  const double *tv = T2(o->tensorvisc, ie);
  gradient_sphere(o, ie, nl, f, b0, b1);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double g0 = AT(b0, ig, jg, k), g1 = AT(b1, ig, jg, k);
        AT(s0, ig, jg, k) = M4(tv, 0, 0, ig, jg) * g0 + M4(tv, 1, 0, ig, jg) * g1;
        AT(s1, ig, jg, k) = M4(tv, 0, 1, ig, jg) * g0 + M4(tv, 1, 1, ig, jg) * g1;
      }
  divergence_sphere_wk(o, ie, nl, s0, s1, lap);
}

static void grad_sphere_wk_testcov(const Oracle *o, int ie, int nl, const double *s, double *g0, double *g1) {
  const double *D = T2(o->D, ie), *mp = E2(o->mp, ie), *mi = T2(o->metinv, ie), *metdet = E2(o->metdet, ie);
  for (int n = 0; n < NP; ++n)
    for (int m = 0; m < NP; ++m)
      for (int k = 0; k < nl; ++k) {
        double b0 = 0, b1 = 0;
        for (int j = 0; j < NP; ++j) {
          const double mpnj = M2(mp, n, j), mpjm = M2(mp, j, m), md = M2(metdet, n, m);
          const double snj = AT(s, n, j, k), sjm = AT(s, j, m, k);
          const double djm = o->dvv[j][m], djn = o->dvv[j][n];
          b0 -= (mpnj * M4(mi, 0, 0, n, m) * md * snj * djm + mpjm * M4(mi, 0, 1, n, m) * md * sjm * djn);
          b1 -= (mpnj * M4(mi, 1, 0, n, m) * md * snj * djm + mpjm * M4(mi, 1, 1, n, m) * md * sjm * djn);
        }
        AT(g0, n, m, k) = (M4(D, 0, 0, n, m) * b0 + M4(D, 1, 0, n, m) * b1) * RREARTH;
        AT(g1, n, m, k) = (M4(D, 0, 1, n, m) * b0 + M4(D, 1, 1, n, m) * b1) * RREARTH;
      }
}

static void curl_sphere_wk_testcov_update(const Oracle *o, int ie, int nl, double alpha, double beta,
                                          const double *s, double *c0, double *c1) {
  const double *D = T2(o->D, ie), *mp = E2(o->mp, ie);
  for (int n = 0; n < NP; ++n)
    for (int m = 0; m < NP; ++m)
      for (int k = 0; k < nl; ++k) {
        double sb0 = 0, sb1 = 0;
        for (int j = 0; j < NP; ++j) {
          sb0 -= M2(mp, j, m) * AT(s, j, m, k) * o->dvv[j][n];
          sb1 += M2(mp, n, j) * AT(s, n, j, k) * o->dvv[j][m];
        }
        AT(c0, n, m, k) = beta * AT(c0, n, m, k) + alpha * (M4(D, 0, 0, n, m) * sb0 + M4(D, 1, 0, n, m) * sb1) * RREARTH;
        AT(c1, n, m, k) = beta * AT(c1, n, m, k) + alpha * (M4(D, 0, 1, n, m) * sb0 + M4(D, 1, 1, n, m) * sb1) * RREARTH;
      }
}

/* vector laplacian, contravariant form.  Output may alias input (per-point
 * update at the end).  scratch: 4 element blocks. */
static void vlaplace_sphere_wk_contra(const Oracle *o, int ie, int nl, double nu_ratio,
                                      const double *u0, const double *u1, double *l0, double *l1,
                                      double *scr) {
  double *div = scr, *gc0 = scr + E3, *gc1 = scr + 2 * E3, *t0 = scr + 3 * E3, *t1 = scr + 4 * E3;
  const double *sph = E2(o->spheremp, ie);
  divergence_sphere(o, ie, nl, u0, u1, div, t0, t1);
  if (nu_ratio > 0 && nu_ratio != 1.0)
    for (int p = 0; p < NP2; ++p)
      for (int k = 0; k < nl; ++k) div[p * L + k] *= nu_ratio;
  grad_sphere_wk_testcov(o, ie, nl, div, gc0, gc1);
  vorticity_sphere(o, ie, nl, u0, u1, div, t0, t1); /* div buffer now holds vorticity */
  curl_sphere_wk_testcov_update(o, ie, nl, -1.0, 1.0, div, gc0, gc1);
  const double re2 = RREARTH * RREARTH;
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double f = 2.0 * M2(sph, ig, jg);
        const double a = AT(u0, ig, jg, k), b = AT(u1, ig, jg, k);
        AT(l0, ig, jg, k) = f * a * re2 + AT(gc0, ig, jg, k);
        AT(l1, ig, jg, k) = f * b * re2 + AT(gc1, ig, jg, k);
      }
}

/* vlaplace_sphere_wk_cartesian (SphereOperators.hpp:750-814, UNDAMPRRCART):
 * tensor laplacian of the three cartesian components, projected back, plus
 * 2 spheremp v / a^2.  Output may alias input.  scratch: 7 element blocks. */
static void vlaplace_sphere_wk_cartesian(const Oracle *o, int ie, int nl, const double *u0, const double *u1,
                                         double *l0, double *l1, double *scr) {
  const double *vs = o->vec_sph2cart + (size_t)ie * 6 * NP2; /* [2][3][NP][NP] */
  const double *sph = E2(o->spheremp, ie);
#define VS(c, i, ig, jg) vs[(((c) * 3 + (i)) * NP + (ig)) * NP + (jg)]
  double *lc[3] = {scr, scr + E3, scr + 2 * E3};
  double *w = scr + 3 * E3;
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double v0 = AT(u0, ig, jg, k), v1 = AT(u1, ig, jg, k);
        for (int i = 0; i < 3; ++i) AT(lc[i], ig, jg, k) = VS(0, i, ig, jg) * v0 + VS(1, i, ig, jg) * v1;
      }
  for (int i = 0; i < 3; ++i) laplace_tensor(o, ie, nl, lc[i], lc[i], w, w + E3, w + 2 * E3, w + 3 * E3);
  for (int ig = 0; ig < NP; ++ig)
    for (int jg = 0; jg < NP; ++jg)
      for (int k = 0; k < nl; ++k) {
        const double v0 = AT(u0, ig, jg, k), v1 = AT(u1, ig, jg, k);
        const double a = AT(lc[0], ig, jg, k), b = AT(lc[1], ig, jg, k), c = AT(lc[2], ig, jg, k);
        const double und = PLANT(11) ? 0.0 : 2.0; /* plant 11: undamping term dropped */
        AT(l0, ig, jg, k) = VS(0, 0, ig, jg) * a + VS(0, 1, ig, jg) * b + VS(0, 2, ig, jg) * c +
                            und * M2(sph, ig, jg) * v0 * (RREARTH) * (RREARTH);
        AT(l1, ig, jg, k) = VS(1, 0, ig, jg) * a + VS(1, 1, ig, jg) * b + VS(1, 2, ig, jg) * c +
                            und * M2(sph, ig, jg) * v1 * (RREARTH) * (RREARTH);
      }
#undef VS
}

/* ==================================================================== DSS
 * Boundary exchange on one rank (BoundaryExchange.cpp:285-618): every
 * element packs its edge/corner values into the neighbour's receive slot,
 * then each element accumulates its slots in the fixed order
 *   for k in 0..3: S, N, W, E   then corners SW, SE, NW, NE
 * and (optionally) multiplies every point by rspheremp. */
static const int EDGE_PTS_FWD[4][NP][2] = {
    {{0, 0}, {0, 1}, {0, 2}, {0, 3}},   /* S */
    {{3, 0}, {3, 1}, {3, 2}, {3, 3}},   /* N */
    {{0, 0}, {1, 0}, {2, 0}, {3, 0}},   /* W */
    {{0, 3}, {1, 3}, {2, 3}, {3, 3}}};  /* E */
static const int CORNER_PT[4][2] = {{0, 0}, {0, 3}, {3, 0}, {3, 3}};
/* direction of an edge connection given (local pos, remote pos): 1 = backward */
static const int EDGE_DIR[4][4] = {{1, 0, 0, 1}, {0, 1, 1, 0}, {0, 1, 1, 0}, {1, 0, 0, 1}};

static void conn_point(int pos, int dir, int k, int *ig, int *jg) {
  if (pos < 4) {
    const int kk = dir ? NP - 1 - k : k;
    *ig = EDGE_PTS_FWD[pos][kk][0];
    *jg = EDGE_PTS_FWD[pos][kk][1];
  } else {
    *ig = CORNER_PT[pos - 4][0];
    *jg = CORNER_PT[pos - 4][1];
  }
}

/* fields[f] = base pointer of field f; element ie's block at fields[f] + ie*estride[f] */
static void dss(Oracle *o, int nf, double **fields, const size_t *estride, int nl, const double *rsphmp) {
  if (nf > o->dss_buf_fields) {
    free(o->dss_buf);
    o->dss_buf = alloc0((size_t)o->nelem * NCONN * NP * nf * L);
    o->dss_buf_fields = nf;
  }
  const size_t slot = (size_t)NP * nf * L; /* per (elem, conn) */
  /* pack */
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int c = 0; c < NCONN; ++c) {
      const int rl = o->conn_lid[ie * NCONN + c];
      if (rl < 0) continue;
      const int rp = o->conn_pos[ie * NCONN + c], dir = o->conn_dir[ie * NCONN + c];
      double *buf = o->dss_buf + ((size_t)rl * NCONN + rp) * slot;
      const int npts = c < 4 ? NP : 1;
      for (int k = 0; k < npts; ++k) {
        int ig, jg;
        conn_point(c, dir, k, &ig, &jg);
        for (int f = 0; f < nf; ++f) {
          const double *src = fields[f] + (size_t)ie * estride[f];
          for (int lv = 0; lv < nl; ++lv) buf[((size_t)k * nf + f) * L + lv] = AT(src, ig, jg, lv);
        }
      }
    }
  /* unpack in the fixed order, then scale */
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int f = 0; f < nf; ++f) {
      double *dst = fields[f] + (size_t)ie * estride[f];
      for (int k = 0; k < NP; ++k)
        for (int e = 0; e < 4; ++e) {
          const double *buf = o->dss_buf + ((size_t)ie * NCONN + e) * slot;
          const int ig = EDGE_PTS_FWD[e][k][0], jg = EDGE_PTS_FWD[e][k][1];
          for (int lv = 0; lv < nl; ++lv) AT(dst, ig, jg, lv) += buf[((size_t)k * nf + f) * L + lv];
        }
      for (int c = 4; c < NCONN; ++c) {
        if (o->conn_lid[ie * NCONN + c] < 0) continue; /* missing corner: nothing to add */
        if (PLANT(6)) continue;                        /* planted: corner copies dropped */
        const double *buf = o->dss_buf + ((size_t)ie * NCONN + c) * slot;
        const int ig = CORNER_PT[c - 4][0], jg = CORNER_PT[c - 4][1];
        for (int lv = 0; lv < nl; ++lv) AT(dst, ig, jg, lv) += buf[(size_t)f * L + lv];
      }
      if (rsphmp) {
        const double *r = E2(rsphmp, ie);
        for (int ig = 0; ig < NP; ++ig)
          for (int jg = 0; jg < NP; ++jg)
            for (int lv = 0; lv < nl; ++lv) AT(dst, ig, jg, lv) *= M2(r, ig, jg);
      }
    }
}

/* element-wise min/max exchange of qlim [ie][q][2][L] (BoundaryExchange.cpp:620-849) */
static void neighbor_minmax(Oracle *o) {
  const size_t es = (size_t)o->qsize * 2 * L;
  double *old = alloc0((size_t)o->nelem * es);
  memcpy(old, o->qlim, sizeof(double) * o->nelem * es);
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int c = 0; c < NCONN; ++c) {
      const int rl = o->conn_lid[ie * NCONN + c];
      if (rl < 0) continue;
      for (int q = 0; q < o->qsize; ++q) {
        double *mine = o->qlim + (size_t)ie * es + (size_t)q * 2 * L;
        const double *nb = old + (size_t)rl * es + (size_t)q * 2 * L;
        for (int k = 0; k < L; ++k) {
          mine[L + k] = mine[L + k] > nb[L + k] ? mine[L + k] : nb[L + k];
          mine[k] = mine[k] < nb[k] ? mine[k] : nb[k];
        }
      }
    }
  free(old);
}

/* =================================================================== CAAR */
static void caar_element(Oracle *o, int ie, int nm1, int n0, int np1, int n0_qdp, double dt,
                         double eta_ave_w, double *w) {
  /* work blocks */
  double *tv = w, *p = w + E3, *divv = w + 2 * E3, *omb = w + 3 * E3, *vdp0 = w + 4 * E3, *vdp1 = w + 5 * E3;
  double *g0 = w + 6 * E3, *g1 = w + 7 * E3, *b0 = w + 8 * E3, *b1 = w + 9 * E3, *ephi = w + 10 * E3;
  double *vort = w + 11 * E3, *eg0 = w + 12 * E3, *eg1 = w + 13 * E3;
  const double *v0n = V(ie, n0, 0), *v1n = V(ie, n0, 1), *tn = T(ie, n0), *dpn = DP(ie, n0);
  double *vn0a = o->vn0 + (size_t)ie * 2 * E3, *vn0b = vn0a + E3;
  double *phi = o->phi + (size_t)ie * E3, *omega = o->omega_p + (size_t)ie * E3;
  const double *sph = E2(o->spheremp, ie), *fcor = E2(o->fcor, ie), *phis = E2(o->phis, ie);
  /* virtual temperature */
  for (size_t i = 0; i < E3; ++i) {
    if (n0_qdp < 0) {
      tv[i] = tn[i];
    } else {
      double qt = QDP(ie, n0_qdp, 0)[i] / dpn[i];
      qt *= (RWATER_VAPOR / RGAS - 1.0);
      qt += 1.0;
      tv[i] = tn[i] * qt;
    }
  }
  /* div(v dp), vn0 accumulation */
  for (size_t i = 0; i < E3; ++i) {
    vdp0[i] = v0n[i] * dpn[i];
    vdp1[i] = v1n[i] * dpn[i];
    vn0a[i] += eta_ave_w * vdp0[i];
    vn0b[i] += eta_ave_w * vdp1[i];
  }
  divergence_sphere(o, ie, L, vdp0, vdp1, divv, b0, b1);
  /* pressure (serial scan, CaarFunctorImpl.hpp:621-650) */
  for (int pt = 0; pt < NP2; ++pt) {
    double dp_prev = 0, p_prev = o->hyai0 * o->ps0;
    for (int k = 0; k < L; ++k) {
      p[pt * L + k] = p_prev + 0.5 * (dp_prev + dpn[pt * L + k]);
      p_prev = p[pt * L + k];
      dp_prev = dpn[pt * L + k];
    }
  }
  /* hydrostatic geopotential, bottom-up (:689-729) */
  for (int pt = 0; pt < NP2; ++pt) {
    double integration = 0;
    for (int k = L - 1; k >= 0; --k) {
      const size_t i = (size_t)pt * L + k;
      const double r = RGAS * tv[i] * (dpn[i] * 0.5 / p[i]);
      phi[i] = phis[pt] + 2.0 * integration + r;
      integration = integration + r;
    }
  }
  /* omega/p (:854-889) */
  gradient_sphere(o, ie, L, p, g0, g1);
  for (int pt = 0; pt < NP2; ++pt) {
    double integration = 0;
    for (int k = 0; k < L; ++k) {
      const size_t i = (size_t)pt * L + k;
      const double vgrad_p = v0n[i] * g0[i] + v1n[i] * g1[i];
      omb[i] = (vgrad_p - (integration + (PLANT(15) ? 1.0 : 0.5) * divv[i])) / p[i]; /* plant 15: the 1/2 dropped */
      integration = integration + divv[i];
    }
  }
  /* phase 3, rsplit == 0 (:136-143, :237-300, :497-597): eta_dot_dpdn from
     the column sums of div(v dp), vertical advection of T and v, and the
     eta_dot_dpdn accumulation; zero vertical flux otherwise */
  double *eta = w + 14 * E3, *tva = w + 15 * E3, *vva0 = w + 16 * E3, *vva1 = w + 17 * E3;
  for (size_t i = 0; i < E3; ++i) eta[i] = tva[i] = vva0[i] = vva1[i] = 0.0;
  if (o->rsplit == 0) {
    for (int pt = 0; pt < NP2; ++pt) {
      double *e = eta + (size_t)pt * L;
      const double *dv = divv + (size_t)pt * L;
      double sdot = 0.0;
      for (int k = 0; k < L - 1; ++k) {
        sdot += dv[k];
        e[k + 1] = sdot;
      }
      sdot += dv[L - 1];
      for (int k = 1; k < L; ++k) e[k] = o->hybi[k] * sdot - e[k];
      e[0] = 0.0;
      const double *dp = dpn + (size_t)pt * L, *t = tn + (size_t)pt * L;
      const double *u = v0n + (size_t)pt * L, *vv = v1n + (size_t)pt * L;
      double *ta = tva + (size_t)pt * L, *ua = vva0 + (size_t)pt * L, *va = vva1 + (size_t)pt * L;
      {
        const double facp = (0.5 * 1 / dp[0]) * e[1];
        ta[0] = facp * (t[1] - t[0]);
        ua[0] = facp * (u[1] - u[0]);
        va[0] = facp * (vv[1] - vv[0]);
      }
      for (int k = 1; k < L - 1; ++k) {
        const double facp = 0.5 * (1 / dp[k]) * e[k + 1];
        const double facm = 0.5 * (1 / dp[k]) * e[k];
        ta[k] = facp * (t[k + 1] - t[k]) + facm * (t[k] - t[k - 1]);
        ua[k] = facp * (u[k + 1] - u[k]) + facm * (u[k] - u[k - 1]);
        va[k] = facp * (vv[k + 1] - vv[k]) + facm * (vv[k] - vv[k - 1]);
      }
      {
        const int k = L - 1;
        const double facm = (0.5 * (1 / dp[k])) * e[k];
        ta[k] = facm * (t[k] - t[k - 1]);
        ua[k] = facm * (u[k] - u[k - 1]);
        va[k] = facm * (vv[k] - vv[k - 1]);
      }
    }
    double *ea = o->eta_dot_dpdn + (size_t)ie * E3;
    for (size_t i = 0; i < E3; ++i) ea[i] += eta_ave_w * eta[i];
  }
  /* omega accumulation */
  for (size_t i = 0; i < E3; ++i) omega[i] += eta_ave_w * omb[i];
  /* temperature */
  gradient_sphere(o, ie, L, tn, b0, b1);
  double *tnp1 = T(ie, np1);
  const double *tnm1 = T(ie, nm1);
  /* velocity needs v(n0) after T(np1) is written when n0 == np1: T(np1) does not alias v */
  for (int pt = 0; pt < NP2; ++pt)
    for (int k = 0; k < L; ++k) {
      const size_t i = (size_t)pt * L + k;
      const double vgrad_t = v0n[i] * b0[i] + v1n[i] * b1[i];
      const double ttens = (o->rsplit > 0 ? 0 : (PLANT(12) ? tva[i] : -tva[i])) - vgrad_t + (PLANT(16) ? -KAPPA : KAPPA) * tv[i] * omb[i]; /* plants 12, 16 */
      double temp = ttens * dt + tnm1[i];
      temp *= sph[pt];
      tnp1[i] = temp;
    }
  /* velocity (:98-132, :184-232) */
  for (size_t i = 0; i < E3; ++i) {
    eg0[i] = RGAS * (tv[i] / p[i]) * g0[i];
    eg1[i] = RGAS * (tv[i] / p[i]) * g1[i];
    if (PLANT(3)) { eg0[i] = -eg0[i]; eg1[i] = -eg1[i]; } /* planted: pressure-gradient sign */
    ephi[i] = 0.5 * (v0n[i] * v0n[i] + v1n[i] * v1n[i]) + phi[i];
    if (PLANT(2)) ephi[i] = phi[i]; /* planted: kinetic-energy gradient dropped */
  }
  gradient_sphere_update(o, ie, ephi, eg0, eg1);
  vorticity_sphere(o, ie, L, v0n, v1n, vort, b0, b1);
  double *v0p = V(ie, np1, 0), *v1p = V(ie, np1, 1);
  const double *v0m = V(ie, nm1, 0), *v1m = V(ie, nm1, 1);
  for (int pt = 0; pt < NP2; ++pt)
    for (int k = 0; k < L; ++k) {
      const size_t i = (size_t)pt * L + k;
      if (PLANT(1)) vort[i] -= fcor[pt]; /* planted: Coriolis sign */
      else vort[i] += fcor[pt];
      eg0[i] *= -1;
      eg0[i] += (o->rsplit > 0 ? 0 : -vva0[i]) + v1n[i] * vort[i];
      eg1[i] *= -1;
      eg1[i] += (o->rsplit > 0 ? 0 : -vva1[i]) - v0n[i] * vort[i];
      eg0[i] *= dt;
      eg0[i] += v0m[i];
      eg1[i] *= dt;
      eg1[i] += v1m[i];
      const double n0v = sph[pt] * eg0[i], n1v = sph[pt] * eg1[i];
      v0p[i] = n0v;
      v1p[i] = n1v;
    }
  /* dp3d (:468-493); the eta_dot_dpdn buffer is zero when rsplit > 0 */
  double *dpp = DP(ie, np1);
  const double *dpm = DP(ie, nm1);
  for (int pt = 0; pt < NP2; ++pt)
    for (int k = 0; k < L; ++k) {
      const size_t i = (size_t)pt * L + k;
      double tmp = k + 1 < L ? eta[i + 1] : 0;
      tmp += divv[i];
      tmp -= eta[i];
      tmp = dpm[i] - tmp * dt;
      dpp[i] = sph[pt] * tmp;
    }
}

#define NWORK 18
static double *work_for_thread(double *work, size_t e3) {
#ifdef _OPENMP
  return work + (size_t)omp_get_thread_num() * NWORK * e3;
#else
  (void)e3;
  return work;
#endif
}

static int nthreads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static void caar_run(Oracle *o, int nm1, int n0, int np1, double dt, double eta_ave_w, double *work) {
  const int n0_qdp = o->moist ? o->n0_qdp : -1;
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie) caar_element(o, ie, nm1, n0, np1, n0_qdp, dt, eta_ave_w, work_for_thread(work, E3));
  double *f[4] = {V(0, np1, 0), V(0, np1, 1), T(0, np1), DP(0, np1)};
  const size_t es[4] = {6 * E3, 6 * E3, 3 * E3, 3 * E3};
  dss(o, 4, f, es, L, o->rspheremp);
}

static int stop_here(Oracle *o, int phase) {
  if (o->debug_stop == phase) o->stopped = 1;
  return o->stopped;
}

static void u3_5stage(Oracle *o, int nm1, int n0, int np1, double dt, double eta_ave_w, double *work) {
  caar_run(o, n0, n0, nm1, dt / 5.0, eta_ave_w / 4.0, work);
  if (stop_here(o, 1)) return;
  caar_run(o, n0, nm1, np1, dt / 5.0, 0.0, work);
  caar_run(o, n0, np1, np1, dt / 3.0, 0.0, work);
  caar_run(o, n0, np1, np1, 2.0 * dt / 3.0, 0.0, work);
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (size_t i = 0; i < E3; ++i) {
#if ORACLE_PLANT == 17 /* plant 17: stage base (4 u1 - u0)/3 */
      T(ie, nm1)[i] = (4.0 * T(ie, nm1)[i] - T(ie, n0)[i]) / 3.0;
      V(ie, nm1, 0)[i] = (4.0 * V(ie, nm1, 0)[i] - V(ie, n0, 0)[i]) / 3.0;
      V(ie, nm1, 1)[i] = (4.0 * V(ie, nm1, 1)[i] - V(ie, n0, 1)[i]) / 3.0;
      DP(ie, nm1)[i] = (4.0 * DP(ie, nm1)[i] - DP(ie, n0)[i]) / 3.0;
#else
      T(ie, nm1)[i] = (5.0 * T(ie, nm1)[i] - T(ie, n0)[i]) / 4.0;
      V(ie, nm1, 0)[i] = (5.0 * V(ie, nm1, 0)[i] - V(ie, n0, 0)[i]) / 4.0;
      V(ie, nm1, 1)[i] = (5.0 * V(ie, nm1, 1)[i] - V(ie, n0, 1)[i]) / 4.0;
      DP(ie, nm1)[i] = (5.0 * DP(ie, nm1)[i] - DP(ie, n0)[i]) / 4.0;
#endif
    }
  caar_run(o, nm1, np1, np1, 3.0 * dt / 4.0, 3.0 * eta_ave_w / 4.0, work);
}

/* ======================================================== hyperviscosity */
static void hv_run(Oracle *o, int np1, double dt_in, double eta_ave_w, double *work) {
  const double dt = dt_in / o->hypervis_subcycle;
  const double nu_scale_top[3] = {(PLANT(18) ? -4.0 : 4.0) * o->nu_top, 2.0 * o->nu_top, 1.0 * o->nu_top}; /* plant 18 */
  for (int cyc = 0; cyc < o->hypervis_subcycle; ++cyc) {
    /* first laplacian */
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie) {
      double *w = work_for_thread(work, E3);
      double *tt = o->ttens + (size_t)ie * E3, *dpt = o->dptens + (size_t)ie * E3;
      double *vt0 = o->vtens + (size_t)ie * 2 * E3, *vt1 = vt0 + E3;
      laplace_simple(o, ie, L, T(ie, np1), tt, w, w + E3);
      laplace_simple(o, ie, L, DP(ie, np1), dpt, w, w + E3);
      vlaplace_sphere_wk_contra(o, ie, L, o->nu_ratio1, V(ie, np1, 0), V(ie, np1, 1), vt0, vt1, w);
    }
    {
      double *f[4] = {o->vtens, o->vtens + E3, o->ttens, o->dptens};
      const size_t es[4] = {2 * E3, 2 * E3, E3, E3};
      dss(o, 4, f, es, L, o->rspheremp);
    }
    /* second laplacian */
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie) {
      double *w = work_for_thread(work, E3);
      double *tt = o->ttens + (size_t)ie * E3, *dpt = o->dptens + (size_t)ie * E3;
      double *vt0 = o->vtens + (size_t)ie * 2 * E3, *vt1 = vt0 + E3;
      if (o->consthv) {
        laplace_simple(o, ie, L, tt, tt, w, w + E3);
        laplace_simple(o, ie, L, dpt, dpt, w, w + E3);
        vlaplace_sphere_wk_contra(o, ie, L, o->nu_ratio2, vt0, vt1, vt0, vt1, w);
      } else {
        /* TagSecondLaplaceTensorHV (HyperviscosityFunctorImpl.hpp:110-128) */
        laplace_tensor(o, ie, L, tt, tt, w, w + E3, w + 2 * E3, w + 3 * E3);
        laplace_tensor(o, ie, L, dpt, dpt, w, w + E3, w + 2 * E3, w + 3 * E3);
        vlaplace_sphere_wk_cartesian(o, ie, L, vt0, vt1, vt0, vt1, w);
      }
    }
    /* pre-exchange */
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie) {
      double *w = work_for_thread(work, E3);
      double *tt = o->ttens + (size_t)ie * E3, *dpt = o->dptens + (size_t)ie * E3;
      double *vt0 = o->vtens + (size_t)ie * 2 * E3, *vt1 = vt0 + E3;
      double *dpa = o->dpdiss_ave + (size_t)ie * E3, *dpb = o->dpdiss_bih + (size_t)ie * E3;
      const double *dpn = DP(ie, np1);
      const double *sph = E2(o->spheremp, ie);
      for (size_t i = 0; i < E3; ++i) {
        dpa[i] += eta_ave_w * dpn[i] / o->hypervis_subcycle;
        dpb[i] += eta_ave_w * dpt[i] / o->hypervis_subcycle;
      }
      double *lv0 = w + 6 * E3, *lv1 = w + 7 * E3, *lt = w + 8 * E3, *ldp = w + 9 * E3;
      const int nt = L < 3 ? L : 3;
      if (o->nu_top > 0) {
        vlaplace_sphere_wk_contra(o, ie, nt, 1.0, V(ie, np1, 0), V(ie, np1, 1), lv0, lv1, w);
        laplace_simple(o, ie, nt, T(ie, np1), lt, w, w + E3);
        laplace_simple(o, ie, nt, DP(ie, np1), ldp, w, w + E3);
      }
      for (int pt = 0; pt < NP2; ++pt) {
        for (int k = 0; k < L; ++k) {
          const size_t i = (size_t)pt * L + k;
          vt0[i] *= -o->nu;
          vt1[i] *= -o->nu;
          tt[i] *= PLANT(4) ? o->nu_s : -o->nu_s; /* plant 4: biharmonic sign */
          dpt[i] *= -o->nu_p;
        }
        if (o->nu_top > 0)
          for (int k = 0; k < nt; ++k) {
            const size_t i = (size_t)pt * L + k;
            vt0[i] += nu_scale_top[k] * lv0[i];
            vt1[i] += nu_scale_top[k] * lv1[i];
            tt[i] += nu_scale_top[k] * lt[i];
            dpt[i] += nu_scale_top[k] * ldp[i];
          }
        for (int k = 0; k < L; ++k) {
          const size_t i = (size_t)pt * L + k;
          dpt[i] *= dt;
          dpt[i] += dpn[i] * sph[pt];
        }
      }
    }
    {
      double *f[4] = {o->vtens, o->vtens + E3, o->ttens, o->dptens};
      const size_t es[4] = {2 * E3, 2 * E3, E3, E3};
      dss(o, 4, f, es, L, NULL);
    }
    /* update states */
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie) {
      double *tt = o->ttens + (size_t)ie * E3, *dpt = o->dptens + (size_t)ie * E3;
      double *vt0 = o->vtens + (size_t)ie * 2 * E3, *vt1 = vt0 + E3;
      double *v0 = V(ie, np1, 0), *v1 = V(ie, np1, 1), *t = T(ie, np1), *dp = DP(ie, np1);
      const double *r = E2(o->rspheremp, ie);
      for (int pt = 0; pt < NP2; ++pt)
        for (int k = 0; k < L; ++k) {
          const size_t i = (size_t)pt * L + k;
          vt0[i] = (dt * vt0[i] * r[pt]);
          vt1[i] = (dt * vt1[i] * r[pt]);
          v0[i] += vt0[i];
          v1[i] += vt1[i];
          tt[i] = (dt * tt[i] * r[pt]);
          const double heating = vt0[i] * v0[i] + vt1[i] * v1[i];
          t[i] = t[i] + tt[i] - heating / CP;
          dp[i] = (dpt[i] * r[pt]);
        }
    }
  }
}

/* ============================================================== limiters */
static void limit8(double mass, double minp, double maxp, double *x, const double *c) {
  const int maxiter = NP2 - 1;
  const double tol_limiter = 5e-14;
  for (int iter = 0; iter < maxiter; ++iter) {
    double addmass = 0;
    for (int k = 0; k < NP2; ++k) {
      double delta = 0;
      if (x[k] > maxp) {
        delta = x[k] - maxp;
        x[k] = maxp;
      } else if (x[k] < minp) {
        delta = x[k] - minp;
        x[k] = minp;
      }
      addmass += delta * c[k];
    }
    if (fabs(addmass) <= tol_limiter * fabs(mass)) break;
    double weightssum = 0;
    if (addmass > 0) {
      for (int k = 0; k < NP2; ++k)
        if (x[k] < maxp) weightssum += c[k];
      const double adw = addmass / weightssum;
      for (int k = 0; k < NP2; ++k) x[k] += (x[k] < maxp) ? adw : 0;
    } else {
      for (int k = 0; k < NP2; ++k)
        if (x[k] > minp) weightssum += c[k];
      const double adw = addmass / weightssum;
      for (int k = 0; k < NP2; ++k) x[k] += (x[k] > minp) ? adw : 0;
    }
  }
}

static void limit9(double mass, double minp, double maxp, double *x, const double *c) {
  (void)mass;
  double addmass = 0;
  for (int k = 0; k < NP2; ++k) {
    double delta = 0;
    if (x[k] > maxp) {
      delta = x[k] - maxp;
      x[k] = maxp;
    } else if (x[k] < minp) {
      delta = x[k] - minp;
      x[k] = minp;
    }
    addmass += delta * c[k];
  }
  if (addmass == 0) return;
  if (addmass > 0) {
    double fac = 0;
    for (int k = 0; k < NP2; ++k) fac += c[k] * (maxp - x[k]);
    if (fac > 0) {
      fac = addmass / fac;
      for (int k = 0; k < NP2; ++k) x[k] += fac * (maxp - x[k]);
    }
  } else {
    double fac = 0;
    for (int k = 0; k < NP2; ++k) fac += c[k] * (x[k] - minp);
    if (fac > 0) {
      fac = addmass / fac;
      for (int k = 0; k < NP2; ++k) x[k] += fac * (x[k] - minp);
    }
  }
}

/* with_limiter_shell (EulerStepFunctorImpl.hpp:693-761), one (element, tracer) */
static void limiter_shell(int option, int nlev, const double *sph, const double *dpmass, double *qlim, double *ptens) {
  for (int k = 0; k < nlev; ++k) {
    double x[NP2], c[NP2];
    for (int p = 0; p < NP2; ++p) {
      const double dpm = dpmass[p * nlev + k];
      c[p] = sph[p] * dpm;
      x[p] = ptens[p * nlev + k] / dpm;
    }
    double s0 = 0, s1 = 0;
    for (int p = 0; p < NP2; ++p) {
      s0 += x[p] * c[p];
      s1 += c[p];
    }
    if (s1 <= 0) continue;
    double minp = qlim[k], maxp = qlim[nlev + k];
    if (minp < 0) minp = qlim[k] = 0;
    if (s0 < minp * s1) minp = qlim[k] = s0 / s1;
    if (s0 > maxp * s1) maxp = qlim[nlev + k] = s0 / s1;
    if (option == 8) limit8(s0, minp, maxp, x, c);
    else limit9(s0, minp, maxp, x, c);
    for (int p = 0; p < NP2; ++p) ptens[p * nlev + k] = x[p] * dpmass[p * nlev + k];
  }
}

/* ============================================================ Euler step */
enum { DSS_ETA = 0, DSS_OMEGA = 1, DSS_DIV_VDP_AVE = 2 };

static void euler_step(Oracle *o, int np1_qdp, int n0_qdp, double dt, double rhs_multiplier, int dssopt,
                       double *rhs_viss, double *work) {
  const int qs = o->qsize;
  const int lim = o->limiter_option;
  /* compute_dp */
  const double rhsmdt = rhs_multiplier * dt;
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (size_t i = 0; i < E3; ++i)
      o->dpbuf[ie * E3 + i] = o->derived_dp[ie * E3 + i] - rhsmdt * o->divdp_proj[ie * E3 + i];
  /* compute_qmin_qmax */
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int q = 0; q < qs; ++q) {
      const double *qd = QDP(ie, n0_qdp, q), *dp = o->dpbuf + ie * E3;
      double *qb = o->qtens_bih + ((size_t)ie * qs + q) * E3;
      double *ql = o->qlim + ((size_t)ie * qs + q) * 2 * L;
      if (rhs_multiplier != 1.0)
        for (int k = 0; k < L; ++k) {
          const double vv = qd[k] / dp[k];
          qb[k] = vv;
          ql[k] = vv;
          ql[L + k] = vv;
        }
      for (int pt = 0; pt < NP2; ++pt)
        for (int k = 0; k < L; ++k) {
          const double vv = qd[pt * L + k] / dp[pt * L + k];
          qb[pt * L + k] = vv;
          ql[k] = ql[k] < vv ? ql[k] : vv;
          ql[L + k] = ql[L + k] > vv ? ql[L + k] : vv;
        }
    }
  if (rhs_multiplier == 0.0) {
    neighbor_minmax(o);
  } else if (rhs_multiplier == 2.0) {
    /* minmax_and_biharmonic: the min/max values are packed before the
       biharmonic and unpacked after it; qlim is not touched in between */
    *rhs_viss = 3.0;
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie) {
      double *w = work_for_thread(work, E3);
      for (int q = 0; q < qs; ++q) {
        double *qb = o->qtens_bih + ((size_t)ie * qs + q) * E3;
        if (o->nu_p > 0) {
          const double *dpa = o->dpdiss_ave + ie * E3;
          for (int pt = 0; pt < NP2; ++pt)
            for (int k = 0; k < L; ++k) qb[pt * L + k] = qb[pt * L + k] * dpa[pt * L + k] / o->dp0[k];
        }
        laplace_simple(o, ie, L, qb, qb, w, w + E3);
      }
    }
    {
      double **f = (double **)malloc(sizeof(double *) * (qs ? qs : 1));
      size_t *es = (size_t *)malloc(sizeof(size_t) * (qs ? qs : 1));
      for (int q = 0; q < qs; ++q) {
        f[q] = o->qtens_bih + (size_t)q * E3;
        es[q] = (size_t)qs * E3;
      }
      dss(o, qs, f, es, L, o->rspheremp);
      free(f);
      free(es);
    }
    const double fac = -(*rhs_viss) * dt * o->nu_q;
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie) {
      double *w = work_for_thread(work, E3);
      const double *sph = E2(o->spheremp, ie);
      for (int q = 0; q < qs; ++q) {
        double *qb = o->qtens_bih + ((size_t)ie * qs + q) * E3;
        if (o->consthv) laplace_simple(o, ie, L, qb, qb, w, w + E3);
        else laplace_tensor(o, ie, L, qb, qb, w, w + E3, w + 2 * E3, w + 3 * E3);
        for (int pt = 0; pt < NP2; ++pt)
          for (int k = 0; k < L; ++k) qb[pt * L + k] = (fac * o->dp0[k] * qb[pt * L + k] / sph[pt]);
      }
    }
    neighbor_minmax(o);
  }
  /* advect_and_limit: setup phase */
  const int add_ps_diss = o->nu_p > 0 && *rhs_viss != 0.0;
  const double diss_fac = add_ps_diss ? -(*rhs_viss) * dt * o->nu_q : 0;
  double *fdss = dssopt == DSS_ETA ? o->eta_dot_dpdn : dssopt == DSS_OMEGA ? o->omega_p : o->divdp_proj;
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie) {
    const double *sph = E2(o->spheremp, ie);
    for (int pt = 0; pt < NP2; ++pt)
      for (int k = 0; k < L; ++k) {
        const size_t i = (size_t)ie * E3 + (size_t)pt * L + k;
        const double dp = o->dpbuf[i];
        o->vstar[(size_t)ie * 2 * E3 + (size_t)pt * L + k] = o->vn0[(size_t)ie * 2 * E3 + (size_t)pt * L + k] / dp;
        o->vstar[(size_t)ie * 2 * E3 + E3 + (size_t)pt * L + k] =
            o->vn0[(size_t)ie * 2 * E3 + E3 + (size_t)pt * L + k] / dp;
        o->dpdissk[i] = dp - dt * o->divdp[i];
        if (add_ps_diss) o->dpdissk[i] += diss_fac * o->dpdiss_bih[i] / sph[pt];
        fdss[i] *= sph[pt];
      }
  }
  /* tracer phase */
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie) {
    double *w = work_for_thread(work, E3);
    const double *sph = E2(o->spheremp, ie);
    for (int q = 0; q < qs; ++q) {
      double *qb = o->qtens_bih + ((size_t)ie * qs + q) * E3;
      divergence_sphere_update(o, ie, -dt, *rhs_viss != 0.0, o->vstar + (size_t)ie * 2 * E3,
                               o->vstar + (size_t)ie * 2 * E3 + E3, QDP(ie, n0_qdp, q), qb, w, w + E3);
      if (lim == 8 || lim == 9)
        limiter_shell(lim, L, sph, o->dpdissk + (size_t)ie * E3, o->qlim + ((size_t)ie * qs + q) * 2 * L, qb);
      double *dst = QDP(ie, np1_qdp, q);
      for (int pt = 0; pt < NP2; ++pt)
        for (int k = 0; k < L; ++k) dst[pt * L + k] = sph[pt] * qb[pt * L + k];
    }
  }
  /* DSS of Qdp(np1) and the extra field */
  {
    double **f = (double **)malloc(sizeof(double *) * (qs + 1));
    size_t *es = (size_t *)malloc(sizeof(size_t) * (qs + 1));
    for (int q = 0; q < qs; ++q) {
      f[q] = QDP(0, np1_qdp, q);
      es[q] = (size_t)2 * qs * E3;
    }
    f[qs] = fdss;
    es[qs] = E3;
    dss(o, qs + 1, f, es, L, o->rspheremp);
    free(f);
    free(es);
  }
}

static void advec_tracers_rk2(Oracle *o, double dt, double *work) {
  double rhs_viss = 0.0;
  /* precompute_divdp */
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie) {
    double *w = work_for_thread(work, E3);
    divergence_sphere(o, ie, L, o->vn0 + (size_t)ie * 2 * E3, o->vn0 + (size_t)ie * 2 * E3 + E3,
                      o->divdp + (size_t)ie * E3, w, w + E3);
    memcpy(o->divdp_proj + (size_t)ie * E3, o->divdp + (size_t)ie * E3, sizeof(double) * E3);
  }
  euler_step(o, o->np1_qdp, o->n0_qdp, dt / 2.0, 0.0, DSS_DIV_VDP_AVE, &rhs_viss, work);
  if (stop_here(o, 4)) return;
  euler_step(o, o->np1_qdp, o->np1_qdp, dt / 2.0, 1.0, DSS_ETA, &rhs_viss, work);
  euler_step(o, o->np1_qdp, o->np1_qdp, dt / 2.0, 2.0, DSS_OMEGA, &rhs_viss, work);
  const double rkstage = 3.0;
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int q = 0; q < o->qsize; ++q) {
      const double *a = QDP(ie, o->n0_qdp, q);
      double *b = QDP(ie, o->np1_qdp, q);
      for (size_t i = 0; i < E3; ++i) b[i] = (a[i] + (rkstage - 1) * b[i]) / rkstage;
    }
}

/* ============================================================ PPM remap
 * One column.  Indexing follows the reference's padded arrays restated with
 * an explicit offset: physical level k (0-based) lives at index k + PAD of
 * dpo/ao (two ghost cells each side), so Fortran level j maps to j-1+PAD. */
#define PAD 2

typedef struct {
  double *dpo, *pio, *pin, *z2, *ppmdx; /* ppmdx[10][L+2] */
  int *kid;
} RemapGrid;

static void ppm_grids(int nlev, const double *src, const double *tgt, RemapGrid *g) {
  double *dpo = g->dpo, *pio = g->pio, *pin = g->pin;
  /* partitions (PpmRemap.hpp:506-597) */
  double acc = 0;
  for (int k = 0; k < nlev; ++k) { pio[k] = acc; acc += src[k]; }
  acc = 0;
  for (int k = 0; k < nlev; ++k) { pin[k] = acc; acc += tgt[k]; }
  pio[nlev] = pio[nlev - 1] + src[nlev - 1];
  pin[nlev] = pin[nlev - 1] + tgt[nlev - 1];
  pio[nlev + 1] = pio[nlev] + 1.0;
  pin[nlev] = pio[nlev];
  for (int k = 0; k < nlev; ++k) dpo[k + PAD] = src[k];
  for (int k = 0; k < 2; ++k) {
    dpo[PAD - 1 - k] = dpo[k + PAD];
    dpo[nlev + PAD + k] = dpo[nlev + PAD - 1 - k];
  }
  /* integral bounds (:600-666) */
  for (int k = 0; k < nlev; ++k) {
    int kk = k + 1;
    while (kk - 1 <= nlev + 1 && pio[kk - 1] <= pin[k + 1]) kk++;
    kk--;
    if (kk == nlev + 1) kk = nlev;
    g->kid[k] = kk - 1;
    g->z2[k] = (pin[k + 1] - (pio[kk - 1] + pio[kk]) * 0.5) / dpo[kk - 1 + PAD];
  }
  /* ppmdx coefficients (:367-413); dx(j + off) with off = PAD - 2 = 0 */
  const int W = nlev + 2;
  double *G = g->ppmdx;
#define DX(i) dpo[(i)]
  for (int j = 0; j < nlev + 2; ++j) {
    G[0 * W + j] = DX(j + 1) / (DX(j) + DX(j + 1) + DX(j + 2));
    G[1 * W + j] = (2.0 * DX(j) + DX(j + 1)) / (DX(j + 1) + DX(j + 2));
    G[2 * W + j] = (DX(j + 1) + 2.0 * DX(j + 2)) / (DX(j) + DX(j + 1));
  }
  for (int j = 0; j < nlev + 1; ++j) {
    G[3 * W + j] = DX(j + 1) / (DX(j + 1) + DX(j + 2));
    G[4 * W + j] = 1.0 / (DX(j) + DX(j + 1) + DX(j + 2) + DX(j + 3));
    G[5 * W + j] = (2.0 * DX(j + 1) * DX(j + 2)) / (DX(j + 1) + DX(j + 2));
    G[6 * W + j] = (DX(j) + DX(j + 1)) / (2.0 * DX(j + 1) + DX(j + 2));
    G[7 * W + j] = (DX(j + 3) + DX(j + 2)) / (2.0 * DX(j + 2) + DX(j + 1));
    G[8 * W + j] = DX(j + 1) * (DX(j) + DX(j + 1)) / (2.0 * DX(j + 1) + DX(j + 2));
    G[9 * W + j] = DX(j + 2) * (DX(j + 2) + DX(j + 3)) / (DX(j + 1) + 2.0 * DX(j + 2));
  }
#undef DX
}

static double dmin3(double a, double b, double c) {
  /* Homme::min(a, b, c) = a < min(b,c) ? a : min(b,c) */
  const double bc = b < c ? b : c;
  return a < bc ? a : bc;
}

static double integrate_parabola(double sq, double lin, double cst, double x1, double x2) {
  return (cst * (x2 - x1) + lin * (x2 * x2 - x1 * x1) / 2.0) + sq * (x2 * x2 * x2 - x1 * x1 * x1) / 3.0;
}

/* remap one column of extensive values ``var`` (stride ``st``) in place */
static void ppm_remap_column(int nlev, int remap_alg, const RemapGrid *g, double *var, size_t st, double *scr) {
  double *ao = scr;                /* nlev + 4 */
  double *mass = ao + nlev + 4;    /* nlev + 2 */
  double *dma = mass + nlev + 2;   /* nlev + 2 */
  double *ai = dma + nlev + 2;     /* nlev + 1 */
  double *cf = ai + nlev + 1;      /* 3 * nlev */
  const double *dpo = g->dpo;
  const int W = nlev + 2;
  const double *G = g->ppmdx;
  for (int k = 0; k < nlev; ++k) ao[k + PAD] = var[k * st] / dpo[k + PAD];
  if (remap_alg == 3) {
    for (int k = 0; k < 2; ++k) {
      ao[PAD - 1 - k] = ao[PAD];
      ao[nlev + PAD + k] = ao[nlev + PAD - 1];
    }
  } else {
    for (int k = 0; k < 2; ++k) {
      ao[PAD - 1 - k] = ao[k + PAD];
      ao[nlev + PAD + k] = ao[nlev + PAD - 1 - k];
    }
  }
  double acc = 0;
  for (int k = 0; k < nlev; ++k) { mass[k + 1] = acc; acc += var[k * st]; }
  mass[nlev + 1] = mass[nlev] + var[(nlev - 1) * st];
  /* compute_ppm (:416-503); cell_means(j + P - 1) is ao[j + 1] with PAD = 2 */
#define A(j) ao[(j) + 1] /* Fortran-style a(j), j = -1 .. nlev+2 */
  for (int j = 0; j < nlev + 2; ++j) {
    if ((A(j + 1) - A(j)) * (A(j) - A(j - 1)) > 0.0) {
      const double da = G[0 * W + j] * (G[1 * W + j] * (A(j + 1) - A(j)) + G[2 * W + j] * (A(j) - A(j - 1)));
      dma[j] = dmin3(fabs(da), 2.0 * fabs(A(j) - A(j - 1)), 2.0 * fabs(A(j + 1) - A(j))) * copysign(1.0, da);
    } else {
      dma[j] = 0.0;
    }
  }
  for (int j = 0; j < nlev + 1; ++j)
    ai[j] = A(j) + G[3 * W + j] * (A(j + 1) - A(j)) +
            G[4 * W + j] * (G[5 * W + j] * (G[6 * W + j] - G[7 * W + j]) * (A(j + 1) - A(j)) -
                            G[8 * W + j] * dma[PLANT(7) ? j : j + 1] + G[9 * W + j] * dma[j]); /* plant 7: index */
  for (int jp = 0; jp < nlev; ++jp) {
    const int j = jp + 1;
    double al = ai[j - 1], ar = ai[j];
    const double a = A(j);
    if ((ar - a) * (a - al) <= 0.) {
      al = a;
      ar = a;
    }
    if ((ar - al) * (a - (al + ar) / 2.0) > (ar - al) * (ar - al) / 6.0) al = 3.0 * a - 2.0 * ar;
    if ((ar - al) * (a - (al + ar) / 2.0) < -(ar - al) * (ar - al) / 6.0) ar = 3.0 * a - 2.0 * al;
    cf[0 * nlev + j - 1] = 1.5 * a - (al + ar) / 4.0;
    cf[1 * nlev + j - 1] = ar - al;
    cf[2 * nlev + j - 1] = 3.0 * (-2.0 * a + (al + ar));
  }
  if (remap_alg == 2) {
    cf[0 * nlev + 0] = A(1);
    cf[0 * nlev + 1] = A(2);
    cf[0 * nlev + nlev - 2] = A(nlev - 1);
    cf[0 * nlev + nlev - 1] = A(nlev);
    cf[1 * nlev + 0] = cf[1 * nlev + 1] = cf[2 * nlev + 0] = cf[2 * nlev + 1] = 0.0;
    cf[1 * nlev + nlev - 2] = cf[1 * nlev + nlev - 1] = cf[2 * nlev + nlev - 2] = cf[2 * nlev + nlev - 1] = 0.0;
  }
#undef A
  /* compute_remap (:288-339, CPU form) */
  double prev = 0;
  for (int k = 0; k < nlev; ++k) {
    const int kk = g->kid[k];
    const double x1 = -0.5, x2 = g->z2[k];
    const double integral = integrate_parabola(cf[2 * nlev + kk], cf[1 * nlev + kk], cf[0 * nlev + kk], x1, x2);
    const double m = mass[kk + 1] + integral * dpo[kk + PAD];
    var[k * st] = k > 0 ? m - prev : m;
    prev = m;
  }
}

static int remap_run(Oracle *o, int np1, int np1_qdp, double dt) {
  int bad = 0;
  const int ns = o->rsplit > 0 ? 3 : 0; /* rsplit == 0: the states stay on their levels */
  const int nf = ns + o->qsize;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int ie = 0; ie < o->nelem; ++ie) {
    double *dpn = DP(ie, np1), *ps = PSV(ie, np1);
    double src[512], tgt[512];
    double dpo[520], pio[520], pin[520], z2[512], ppmdx[10 * 520], scr[8 * 520];
    int kid[512];
    RemapGrid g = {dpo, pio, pin, z2, ppmdx, kid};
    for (int pt = 0; pt < NP2; ++pt) {
      double s = 0.0;
      for (int k = 0; k < L; ++k) s += dpn[pt * L + k];
      s += o->hyai0 * o->ps0;
      ps[pt] = s;
    }
    for (int pt = 0; pt < NP2; ++pt) {
      for (int k = 0; k < L; ++k) {
        tgt[k] = (o->hyai[k + 1] - o->hyai[k]) * o->ps0 + (o->hybi[k + 1] - o->hybi[k]) * ps[pt];
      }
      for (int k = 0; k < L; ++k) {
        if (ns) {
          src[k] = dpn[pt * L + k];
        } else {
          /* _RemapFunctorRSplit<false>::compute_source_thickness (RemapFunctor.hpp:62-90) */
          const double *e = o->eta_dot_dpdn + (size_t)ie * E3 + (size_t)pt * L;
          const double en = k + 1 < L ? e[k + 1] : 0;
          const double delta = en - e[k];
          src[k] = tgt[k] + dt * delta;
        }
        if (isnan(src[k]) || src[k] < 0.0) bad |= 1;
      }
      if (bad) continue;
      /* extrinsic states */
      for (int k = 0; k < L && ns; ++k) {
        V(ie, np1, 0)[pt * L + k] *= src[k];
        V(ie, np1, 1)[pt * L + k] *= src[k];
        T(ie, np1)[pt * L + k] *= src[k];
      }
      ppm_grids(L, src, tgt, &g);
      for (int f = 0; f < nf; ++f) {
        const int fs = f + 3 - ns;
        double *col = fs == 0 ? V(ie, np1, 0) : fs == 1 ? V(ie, np1, 1) : fs == 2 ? T(ie, np1) : QDP(ie, np1_qdp, fs - 3);
        ppm_remap_column(L, o->remap_alg, &g, col + pt * L, 1, scr);
      }
      for (int k = 0; k < L && ns; ++k) {
        V(ie, np1, 0)[pt * L + k] /= tgt[k];
        V(ie, np1, 1)[pt * L + k] /= tgt[k];
        T(ie, np1)[pt * L + k] /= tgt[k];
      }
    }
  }
  return bad;
}

/* ================================================================ driver */
static void prim_step(Oracle *o, double dt, double *work) {
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (size_t i = 0; i < E3; ++i) {
      o->eta_dot_dpdn[ie * E3 + i] = 0;
      o->vn0[(size_t)ie * 2 * E3 + i] = 0;
      o->vn0[(size_t)ie * 2 * E3 + E3 + i] = 0;
      o->omega_p[ie * E3 + i] = 0;
      if (o->nu_p > 0) {
        o->dpdiss_ave[ie * E3 + i] = 0;
        o->dpdiss_bih[ie * E3 + i] = 0;
      }
      o->derived_dp[ie * E3 + i] = DP(ie, o->n0)[i];
    }
  /* prim_advance_exp with qsplit steps */
  for (int n = 0; n < o->qsplit; ++n) {
    if (n > 0) { /* LEAPFROG */
      const int tmp = o->np1;
      o->np1 = o->nm1; o->nm1 = o->n0; o->n0 = tmp; ++o->nstep;
    }
    if (o->moist) {
      const int it = o->nstep / o->qsplit;
      o->n0_qdp = it % 2 == 0 ? 0 : 1;
      o->np1_qdp = 1 - o->n0_qdp;
    }
    const double eta_ave_w = 1.0 / o->qsplit;
    u3_5stage(o, o->nm1, o->n0, o->np1, dt, eta_ave_w, work);
    if (o->stopped || stop_here(o, 2)) return;
    hv_run(o, o->np1, dt, eta_ave_w, work);
    if (stop_here(o, 3)) return;
  }
  if (o->qsize > 0) {
    const int it = o->nstep / o->qsplit;
    o->n0_qdp = it % 2 == 0 ? 0 : 1;
    o->np1_qdp = 1 - o->n0_qdp;
    advec_tracers_rk2(o, dt * o->qsplit, work);
  }
  stop_here(o, 5);
}

/* ============================================================ public API */
Oracle *oracle_create(int nelem, int nlev, int qsize, int qsize_d) {
  Oracle *o = (Oracle *)calloc(1, sizeof(Oracle));
  o->nelem = nelem;
  o->nlev = nlev;
  o->qsize = qsize;
  o->qsize_d = qsize_d < qsize ? qsize : qsize_d;
  o->last_time_step = 1 << 30;
  o->disable_diagnostics = 1;
  const size_t e3 = (size_t)NP2 * nlev, ne = nelem;
  o->D = alloc0(ne * 64); o->Dinv = alloc0(ne * 64); o->metinv = alloc0(ne * 64);
  o->tensorvisc = alloc0(ne * 64); o->vec_sph2cart = alloc0(ne * 96);
  o->fcor = alloc0(ne * 16); o->mp = alloc0(ne * 16); o->spheremp = alloc0(ne * 16);
  o->rspheremp = alloc0(ne * 16); o->metdet = alloc0(ne * 16); o->phis = alloc0(ne * 16);
  o->hyai = alloc0(nlev + 1); o->hybi = alloc0(nlev + 1);
  o->hyai_delta = alloc0(nlev); o->hybi_delta = alloc0(nlev); o->dp0 = alloc0(nlev);
  o->conn_lid = (int *)malloc(sizeof(int) * ne * NCONN);
  o->conn_pos = (int *)malloc(sizeof(int) * ne * NCONN);
  o->conn_dir = (int *)malloc(sizeof(int) * ne * NCONN);
  for (size_t i = 0; i < ne * NCONN; ++i) { o->conn_lid[i] = -1; o->conn_pos[i] = -1; o->conn_dir[i] = 0; }
  o->v = alloc0(ne * 6 * e3); o->t = alloc0(ne * 3 * e3); o->dp3d = alloc0(ne * 3 * e3);
  o->ps_v = alloc0(ne * 3 * 16);
  o->qdp = alloc0(ne * 2 * (qsize ? qsize : 1) * e3);
  o->Q = alloc0(ne * (qsize ? qsize : 1) * e3);
  o->phi = alloc0(ne * e3); o->omega_p = alloc0(ne * e3); o->eta_dot_dpdn = alloc0(ne * e3);
  o->dpdiss_ave = alloc0(ne * e3); o->dpdiss_bih = alloc0(ne * e3); o->derived_dp = alloc0(ne * e3);
  o->divdp = alloc0(ne * e3); o->divdp_proj = alloc0(ne * e3); o->vn0 = alloc0(ne * 2 * e3);
  o->qtens_bih = alloc0(ne * (qsize ? qsize : 1) * e3);
  o->fm = alloc0(ne * 2 * e3); o->ft = alloc0(ne * e3); o->fq = alloc0(ne * (qsize ? qsize : 1) * e3);
  o->qlim = alloc0(ne * (qsize ? qsize : 1) * 2 * nlev);
  o->vstar = alloc0(ne * 2 * e3); o->dpdissk = alloc0(ne * e3); o->dpbuf = alloc0(ne * e3);
  o->ttens = alloc0(ne * e3); o->dptens = alloc0(ne * e3); o->vtens = alloc0(ne * 2 * e3);
  o->consthv = 1;
  return o;
}

void oracle_destroy(Oracle *o) {
  if (!o) return;
  double *ptrs[] = {o->D, o->Dinv, o->metinv, o->tensorvisc, o->vec_sph2cart, o->fcor, o->mp, o->spheremp,
                    o->rspheremp, o->metdet, o->phis, o->hyai, o->hybi, o->hyai_delta, o->hybi_delta, o->dp0,
                    o->v, o->t, o->dp3d, o->ps_v, o->qdp, o->Q, o->phi, o->omega_p, o->eta_dot_dpdn,
                    o->dpdiss_ave, o->dpdiss_bih, o->derived_dp, o->divdp, o->divdp_proj, o->vn0, o->qtens_bih,
                    o->qlim, o->vstar, o->dpdissk, o->dpbuf, o->ttens, o->dptens, o->vtens, o->dss_buf, o->fm, o->ft, o->fq};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  free(o->conn_lid); free(o->conn_pos); free(o->conn_dir);
  free(o);
}

/* same argument meaning as init_simulation_params_c (cxx_f90_interface.cpp:33-117) */
int oracle_set_params(Oracle *o, int remap_alg, int limiter_option, int rsplit, int qsplit, int time_step_type,
                      int qsize, int state_frequency, double nu, double nu_p, double nu_q, double nu_s,
                      double nu_div, double nu_top, int hypervis_order, int hypervis_subcycle,
                      double hypervis_scaling, int ftype, int moist) {
  if (remap_alg < 1 || remap_alg > 3 || (limiter_option != 8 && limiter_option != 9) || time_step_type != 5 ||
      hypervis_order != 2 || rsplit < 0 || qsplit < 1 || !(nu > 0) || !(nu_p > 0) || !(nu_div > 0) ||
      qsize != o->qsize || (ftype != -1 && ftype != 0 && ftype != 2))
    return 1;
  o->remap_alg = remap_alg; o->limiter_option = limiter_option; o->rsplit = rsplit; o->qsplit = qsplit;
  o->time_step_type = time_step_type; o->state_frequency = state_frequency;
  o->nu = nu; o->nu_p = nu_p; o->nu_q = nu_q; o->nu_s = nu_s; o->nu_div = nu_div; o->nu_top = nu_top;
  o->hypervis_order = hypervis_order; o->hypervis_subcycle = hypervis_subcycle;
  o->hypervis_scaling = hypervis_scaling; o->ftype = ftype; o->moist = moist;
  o->consthv = hypervis_scaling == 0.0;
  if (nu != nu_div) {
    const double ratio = nu_div / nu;
    if (hypervis_scaling != 0.0) { o->nu_ratio1 = ratio * ratio; o->nu_ratio2 = 1.0; }
    else { o->nu_ratio1 = ratio; o->nu_ratio2 = ratio; }
  } else {
    o->nu_ratio1 = 1.0; o->nu_ratio2 = 1.0;
  }
  return 0;
}

void oracle_set_dvv(Oracle *o, const double *dvv_f90) {
  for (int a = 0; a < NP; ++a)
    for (int b = 0; b < NP; ++b) o->dvv[a][b] = dvv_f90[a * NP + b];
}

void oracle_set_2d(Oracle *o, const double *D, const double *Dinv, const double *fcor, const double *mp,
                   const double *spheremp, const double *rspheremp, const double *metdet, const double *metinv,
                   const double *phis, const double *tensorvisc, const double *vec_sph2cart) {
  const size_t ne = o->nelem;
  memcpy(o->D, D, sizeof(double) * ne * 64);
  memcpy(o->Dinv, Dinv, sizeof(double) * ne * 64);
  memcpy(o->metinv, metinv, sizeof(double) * ne * 64);
  memcpy(o->fcor, fcor, sizeof(double) * ne * 16);
  memcpy(o->mp, mp, sizeof(double) * ne * 16);
  memcpy(o->spheremp, spheremp, sizeof(double) * ne * 16);
  memcpy(o->rspheremp, rspheremp, sizeof(double) * ne * 16);
  memcpy(o->metdet, metdet, sizeof(double) * ne * 16);
  memcpy(o->phis, phis, sizeof(double) * ne * 16);
  if (tensorvisc) memcpy(o->tensorvisc, tensorvisc, sizeof(double) * ne * 64);
  if (vec_sph2cart) memcpy(o->vec_sph2cart, vec_sph2cart, sizeof(double) * ne * 96);
}

void oracle_set_hvcoord(Oracle *o, double ps0, const double *hyai, const double *hybi) {
  o->ps0 = ps0;
  memcpy(o->hyai, hyai, sizeof(double) * (L + 1));
  memcpy(o->hybi, hybi, sizeof(double) * (L + 1));
  o->hyai0 = hyai[0];
  for (int k = 0; k < L; ++k) {
    o->hyai_delta[k] = hyai[k + 1] - hyai[k];
    o->hybi_delta[k] = hybi[k + 1] - hybi[k];
    o->dp0[k] = o->hyai_delta[k] * ps0 + o->hybi_delta[k] * ps0;
  }
}

/* add_connection semantics (mpi_cxx_f90_interface.cpp:34-55, Connectivity.cpp:77-130):
   1-based Fortran ids and positions W,E,S,N,SW,SE,NW,NE. */
int oracle_add_connection(Oracle *o, int lid1, int gid1, int pos1, int pid1, int lid2, int gid2, int pos2, int pid2) {
  (void)gid1; (void)gid2;
  if (lid1 <= 0 || pos1 <= 0 || lid2 <= 0 || pos2 <= 0 || pid1 <= 0 || pid2 <= 0) return 1;
  if (pid1 != 1 || pid2 != 1) return 2; /* one rank */
  const int p1 = pos1 <= 4 ? ((pos1 - 1) + 2) % 4 : pos1 - 1;
  const int p2 = pos2 <= 4 ? ((pos2 - 1) + 2) % 4 : pos2 - 1;
  if ((p1 < 4) != (p2 < 4)) return 3;
  const int i = (lid1 - 1) * NCONN + p1;
  o->conn_lid[i] = lid2 - 1;
  o->conn_pos[i] = p2;
  o->conn_dir[i] = p1 < 4 ? EDGE_DIR[p1][p2] : 0;
  return 0;
}

/* Fortran (np,np,nlev) block <-> oracle [igp][jgp][lev] */
static void f2o3(const double *f, double *c, int nlev) {
  for (int k = 0; k < nlev; ++k)
    for (int ig = 0; ig < NP; ++ig)
      for (int jg = 0; jg < NP; ++jg) c[(ig * NP + jg) * nlev + k] = f[(k * NP + ig) * NP + jg];
}
static void o2f3(const double *c, double *f, int nlev) {
  for (int k = 0; k < nlev; ++k)
    for (int ig = 0; ig < NP; ++ig)
      for (int jg = 0; jg < NP; ++jg) f[(k * NP + ig) * NP + jg] = c[(ig * NP + jg) * nlev + k];
}

void oracle_set_states(Oracle *o, const double *v, const double *t, const double *dp3d, const double *qdp,
                       const double *ps_v) {
  const size_t e3 = E3;
  for (int ie = 0; ie < o->nelem; ++ie) {
    for (int tl = 0; tl < 3; ++tl) {
      for (int c = 0; c < 2; ++c)
        for (int k = 0; k < L; ++k)
          for (int p = 0; p < NP2; ++p)
            V(ie, tl, c)[p * L + k] = v[((((size_t)ie * 3 + tl) * L + k) * 2 + c) * NP2 + p];
      f2o3(t + ((size_t)ie * 3 + tl) * e3, T(ie, tl), L);
      f2o3(dp3d + ((size_t)ie * 3 + tl) * e3, DP(ie, tl), L);
      memcpy(PSV(ie, tl), ps_v + ((size_t)ie * 3 + tl) * NP2, sizeof(double) * NP2);
    }
    for (int tl = 0; tl < 2; ++tl)
      for (int q = 0; q < o->qsize; ++q) f2o3(qdp + (((size_t)ie * 2 + tl) * o->qsize_d + q) * e3, QDP(ie, tl, q), L);
  }
}

/* f90_push_forcing_to_cxx (cxx_f90_interface.cpp:180-205): FM (np,np,2,nlev),
   FT (np,np,nlev), FQ (np,np,nlev,qsize_d) per element, Fortran order */
void oracle_set_forcing(Oracle *o, const double *fm, const double *ft, const double *fq) {
  const size_t e3 = E3;
  for (int ie = 0; ie < o->nelem; ++ie) {
    for (int c = 0; c < 2; ++c)
      for (int k = 0; k < L; ++k)
        for (int p = 0; p < NP2; ++p)
          o->fm[((size_t)ie * 2 + c) * e3 + p * L + k] = fm[(((size_t)ie * L + k) * 2 + c) * NP2 + p];
    f2o3(ft + (size_t)ie * e3, o->ft + (size_t)ie * e3, L);
    if (fq)
      for (int q = 0; q < o->qsize; ++q)
        f2o3(fq + ((size_t)ie * o->qsize_d + q) * e3, o->fq + ((size_t)ie * o->qsize + q) * e3, L);
  }
}

/* apply_cam_forcing / apply_cam_forcing_dynamics (CamForcing.cpp:20-176):
   T, v at n0 += dt F; for ftype 0 also the tracers at n0_qdp, with the
   negative-mass clip, the moist surface-pressure update from tracer 0
   (summed over levels in order) and Q = Qdp / dp(ps_v). */
static void cam_forcing(Oracle *o, double dt, int tracers) {
  const size_t e3 = E3;
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie) {
    double *t = T(ie, o->n0), *v0 = V(ie, o->n0, 0), *v1 = V(ie, o->n0, 1);
    const double *ft = o->ft + (size_t)ie * e3, *fm0 = o->fm + (size_t)ie * 2 * e3, *fm1 = fm0 + e3;
    for (size_t i = 0; i < e3; ++i) t[i] += (PLANT(9) ? -dt : dt) * ft[i]; /* plant 9: forcing sign */
    for (size_t i = 0; i < e3; ++i) v0[i] += dt * fm0[i];
    for (size_t i = 0; i < e3; ++i) v1[i] += dt * fm1[i];
    if (!tracers) continue;
    if (o->moist && o->qsize > 0) {
      const double *fq0 = o->fq + (size_t)ie * o->qsize * e3, *q0 = QDP(ie, o->n0_qdp, 0);
      for (int p = 0; p < NP2; ++p) {
        double acc = 0.0;
        for (int k = 0; k < L; ++k) {
          double v = dt * fq0[p * L + k];
          const double qs = q0[p * L + k];
          if (qs + v < 0.0 && v < 0.0) v = qs < 0.0 ? 0.0 : -qs;
          acc += v;
        }
        PSV(ie, o->n0)[p] += acc;
      }
    }
    for (int q = 0; q < o->qsize; ++q) {
      double *qd = QDP(ie, o->n0_qdp, q);
      const double *fq = o->fq + ((size_t)ie * o->qsize + q) * e3;
      for (size_t i = 0; i < e3; ++i) {
        double v = dt * fq[i];
        if (!PLANT(10) && qd[i] + v < 0.0 && v < 0.0) v = qd[i] < 0.0 ? 0.0 : -qd[i]; /* plant 10: no clip */
        qd[i] += v;
      }
    }
    for (int q = 0; q < o->qsize; ++q)
      for (int p = 0; p < NP2; ++p)
        for (int k = 0; k < L; ++k) {
          const double dp = o->hyai_delta[k] * o->ps0 + o->hybi_delta[k] * PSV(ie, o->n0)[p];
          o->Q[((size_t)ie * o->qsize + q) * e3 + p * L + k] = QDP(ie, o->n0_qdp, q)[p * L + k] / dp;
        }
  }
}

void oracle_get_states(Oracle *o, double *v, double *t, double *dp3d, double *qdp, double *Q, double *ps_v,
                       double *omega_p) {
  const size_t e3 = E3;
  for (int ie = 0; ie < o->nelem; ++ie) {
    for (int tl = 0; tl < 3; ++tl) {
      if (v)
        for (int c = 0; c < 2; ++c)
          for (int k = 0; k < L; ++k)
            for (int p = 0; p < NP2; ++p)
              v[((((size_t)ie * 3 + tl) * L + k) * 2 + c) * NP2 + p] = V(ie, tl, c)[p * L + k];
      if (t) o2f3(T(ie, tl), t + ((size_t)ie * 3 + tl) * e3, L);
      if (dp3d) o2f3(DP(ie, tl), dp3d + ((size_t)ie * 3 + tl) * e3, L);
      if (ps_v) memcpy(ps_v + ((size_t)ie * 3 + tl) * NP2, PSV(ie, tl), sizeof(double) * NP2);
    }
    for (int q = 0; q < o->qsize; ++q) {
      if (qdp)
        for (int tl = 0; tl < 2; ++tl) o2f3(QDP(ie, tl, q), qdp + (((size_t)ie * 2 + tl) * o->qsize_d + q) * e3, L);
      if (Q) o2f3(o->Q + ((size_t)ie * o->qsize + q) * e3, Q + ((size_t)ie * o->qsize_d + q) * e3, L);
    }
    if (omega_p) o2f3(o->omega_p + (size_t)ie * e3, omega_p + (size_t)ie * e3, L);
  }
}

void oracle_set_time_level(Oracle *o, int nm1, int n0, int np1, int nstep, int nstep0) {
  o->nm1 = nm1 - 1; o->n0 = n0 - 1; o->np1 = np1 - 1; o->nstep = nstep; o->nstep0 = nstep0;
}

/* ------------------------------------------------------------ diagnostics
 * Diagnostics::init (Diagnostics.cpp:17-35): the Fortran arrays, C views
 *   q        [ie][QSIZE_D][lev][igp][jgp]   qvar, qmass [ie][4][QSIZE_D][igp][jgp]
 *   q1mass   [ie][QSIZE_D][igp][jgp]        iener, kener, pener [ie][4][igp][jgp]
 *   iener_wet [ie][igp][jgp]
 * disable_diagnostics as the C-ABI passes it; last_time_step as
 * prim_run_subcycle_c receives it (oracle_set_last_time_step). */
void oracle_set_diagnostics(Oracle *o, int disable_diagnostics, int use_cpstar, double *q, double *qvar,
                            double *qmass, double *q1mass, double *iener, double *iener_wet, double *kener,
                            double *pener) {
  o->disable_diagnostics = disable_diagnostics;
  o->use_cpstar = use_cpstar;
  o->diag_on = !disable_diagnostics && q && qvar && qmass && q1mass && iener && iener_wet && kener && pener;
  o->dg_q = q; o->dg_qvar = qvar; o->dg_qmass = qmass; o->dg_q1mass = q1mass;
  o->dg_iener = iener; o->dg_iener_wet = iener_wet; o->dg_kener = kener; o->dg_pener = pener;
}

void oracle_set_last_time_step(Oracle *o, int last_time_step) { o->last_time_step = last_time_step; }

/* Diagnostics::prim_diag_scalars (Diagnostics.cpp:37-88): the device Q goes to
   the Fortran Q array, then per column and tracer the level sums (in level
   order) of Qdp*Q and of Qdp at the tracer level n0_qdp (before the advance)
   or np1_qdp (after). */
static void prim_diag_scalars(Oracle *o, int before_advance, int ivar) {
  const size_t e3 = E3;
  const int qd = o->qsize_d, t2 = before_advance ? o->n0_qdp : o->np1_qdp;
  for (int ie = 0; ie < o->nelem; ++ie) {
    for (int q = 0; q < qd; ++q) {
      double *f = o->dg_q + ((size_t)ie * qd + q) * e3;
      if (q < o->qsize) o2f3(o->Q + ((size_t)ie * o->qsize + q) * e3, f, L);
      else memset(f, 0, sizeof(double) * e3);
    }
    for (int q = 0; q < o->qsize; ++q)
      for (int p = 0; p < NP2; ++p) {
        const double *qdp = QDP(ie, t2, q) + (size_t)p * L;
        const double *Qf = o->dg_q + ((size_t)ie * qd + q) * e3 + p;
        double acc_qdp_q = 0, acc_qdp = 0;
        for (int k = 0; k < L; ++k) {
          acc_qdp_q += qdp[k] * Qf[(size_t)k * NP2];
          acc_qdp += qdp[k];
        }
        o->dg_qvar[(((size_t)ie * 4 + ivar) * qd + q) * NP2 + p] = acc_qdp_q;
        o->dg_qmass[(((size_t)ie * 4 + ivar) * qd + q) * NP2 + p] = acc_qdp;
        o->dg_q1mass[((size_t)ie * qd + q) * NP2 + p] = acc_qdp;
      }
  }
}

/* Diagnostics::prim_energy_halftimes (Diagnostics.cpp:90-185): per column,
   level sums of cp* T dp, (cp* - cp) T dp, (u^2 + v^2)/2 dp and phis dp at
   the dynamics level n0 (before) or np1 (after), dp from ps_v of that level.
   cp* = cp (1 + (Cpv/cp - 1) Qdp_1/dp) with use_cpstar, else cp.  std::pow(u,2)
   is u*u (both are the correctly rounded square). */
static void prim_energy_halftimes(Oracle *o, int before_advance, int ivar) {
  const int t1 = before_advance ? o->n0 : o->np1, t1q = before_advance ? o->n0_qdp : o->np1_qdp;
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int p = 0; p < NP2; ++p) {
      const double *u = V(ie, t1, 0) + (size_t)p * L, *v = V(ie, t1, 1) + (size_t)p * L;
      const double *t = T(ie, t1) + (size_t)p * L;
      const double ps = PSV(ie, t1)[p], phis = o->phis[(size_t)ie * NP2 + p];
      double ie_ = 0, iew = 0, ke = 0, pe = 0;
      for (int k = 0; k < L; ++k) {
        const double dpt1 = o->hyai_delta[k] * o->ps0 + o->hybi_delta[k] * ps;
        double cps = CP;
        if (o->use_cpstar) {
          const double qv = QDP(ie, t1q, 0)[(size_t)p * L + k] / dpt1;
          cps = CP * (1.0 + (CPWATER_VAPOR / CP - 1.0) * qv);
        }
        ie_ += cps * t[k] * dpt1;
        iew += (cps - CP) * t[k] * dpt1;
        ke += (u[k] * u[k] + v[k] * v[k]) * 0.5 * dpt1;
        pe += (PLANT(13) ? 2.0 : 1.0) * phis * dpt1; /* plant 13: PEner doubled */
      }
      o->dg_iener[((size_t)ie * 4 + ivar) * NP2 + p] = ie_;
      o->dg_iener_wet[(size_t)ie * NP2 + p] = iew;
      o->dg_kener[((size_t)ie * 4 + ivar) * NP2 + p] = ke;
      o->dg_pener[((size_t)ie * 4 + ivar) * NP2 + p] = pe;
    }
}

/* prim_run_subcycle_c (prim_driver.cpp:31-156).  Returns 0, or a nonzero
   error code (1: negative or NaN layer thickness at the remap). */
int oracle_run_subcycle(Oracle *o, double dt, int *nstep, int *nm1, int *n0, int *np1) {
  o->error = 0;
  double *work = alloc0((size_t)nthreads() * NWORK * E3);
  {
    const int it = o->nstep / o->qsplit;
    o->n0_qdp = it % 2 == 0 ? 0 : 1;
    o->np1_qdp = 1 - o->n0_qdp;
  }
  const double dt_q = dt * o->qsplit;
  const double dt_remap = o->rsplit > 0 ? dt_q * o->rsplit : dt_q;
  /* prim_driver.cpp:42-63: diagnostics at the state frequency, the first and
     the last step, unless disabled */
  const int nstep_end = o->nstep + o->qsplit * (o->rsplit > 0 ? o->rsplit : 1);
  const int diag = o->diag_on && ((o->state_frequency > 0 && nstep_end % o->state_frequency == 0) ||
                                  nstep_end == o->nstep0 || nstep_end >= o->last_time_step);
  if (diag) {
    prim_diag_scalars(o, 1, PLANT(14) ? 2 : 3); /* plant 14: slot of the first call */
    prim_energy_halftimes(o, 1, 2);
  }
  if (o->ftype == 0) cam_forcing(o, dt_remap, 1);
  else if (o->ftype == 2) cam_forcing(o, dt_remap, 0);
  if (diag) {
    prim_energy_halftimes(o, 1, 0);
    prim_diag_scalars(o, 1, 0);
  }
  /* dp3d from ps */
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < o->nelem; ++ie)
    for (int p = 0; p < NP2; ++p)
      for (int k = 0; k < L; ++k)
        DP(ie, o->n0)[p * L + k] = o->hyai_delta[k] * o->ps0 + o->hybi_delta[k] * PSV(ie, o->n0)[p];
  o->stopped = 0;
  prim_step(o, dt, work);
  if (o->stopped) {
    free(work);
    *nstep = o->nstep; *nm1 = o->nm1; *n0 = o->n0; *np1 = o->np1;
    return 0;
  }
  for (int r = 1; r < o->rsplit; ++r) {
    const int tmp = o->np1;
    o->np1 = o->nm1; o->nm1 = o->n0; o->n0 = tmp; ++o->nstep;
    prim_step(o, dt, work);
  }
  {
    const int it = o->nstep / o->qsplit;
    o->n0_qdp = it % 2 == 0 ? 0 : 1;
    o->np1_qdp = 1 - o->n0_qdp;
  }
  int rc = remap_run(o, o->np1, o->np1_qdp, dt_remap) ? 1 : 0;
  if (!rc) {
    /* update_q */
#pragma omp parallel for schedule(static)
    for (int ie = 0; ie < o->nelem; ++ie)
      for (int q = 0; q < o->qsize; ++q)
        for (int p = 0; p < NP2; ++p)
          for (int k = 0; k < L; ++k) {
            const double dp = o->hyai_delta[k] * o->ps0 + o->hybi_delta[k] * PSV(ie, o->np1)[p];
            o->Q[((size_t)ie * o->qsize + q) * E3 + p * L + k] = QDP(ie, o->np1_qdp, q)[p * L + k] / dp;
          }
    if (diag) {
      prim_diag_scalars(o, 0, 1);
      prim_energy_halftimes(o, 0, 1);
    }
  }
  {
    const int tmp = o->np1;
    o->np1 = o->nm1; o->nm1 = o->n0; o->n0 = tmp; ++o->nstep;
  }
  *nstep = o->nstep; *nm1 = o->nm1; *n0 = o->n0; *np1 = o->np1;
  free(work);
  if (o->error) rc = o->error;
  return rc;
}

/* -------------------------------------------- unit-level test entry points */
/* single element, nl levels, oracle layout [igp][jgp][lev] with L = nl */
static Oracle *unit_oracle(int nlev, const double *dvv_f90, const double *D, const double *Dinv, const double *metdet,
                           const double *spheremp, const double *mp, const double *metinv) {
  Oracle *o = oracle_create(1, nlev, 0, 0);
  oracle_set_dvv(o, dvv_f90);
  if (D) memcpy(o->D, D, sizeof(double) * 64);
  if (Dinv) memcpy(o->Dinv, Dinv, sizeof(double) * 64);
  if (metdet) memcpy(o->metdet, metdet, sizeof(double) * 16);
  if (spheremp) memcpy(o->spheremp, spheremp, sizeof(double) * 16);
  if (mp) memcpy(o->mp, mp, sizeof(double) * 16);
  if (metinv) memcpy(o->metinv, metinv, sizeof(double) * 64);
  return o;
}

void oracle_unit_gradient(int nlev, const double *dvv, const double *Dinv, const double *s, double *g) {
  Oracle *o = unit_oracle(nlev, dvv, NULL, Dinv, NULL, NULL, NULL, NULL);
  gradient_sphere(o, 0, nlev, s, g, g + NP2 * nlev);
  oracle_destroy(o);
}

void oracle_unit_divergence(int nlev, const double *dvv, const double *Dinv, const double *metdet, const double *v,
                            double *div) {
  Oracle *o = unit_oracle(nlev, dvv, NULL, Dinv, metdet, NULL, NULL, NULL);
  double *b = alloc0(2 * NP2 * nlev);
  divergence_sphere(o, 0, nlev, v, v + NP2 * nlev, div, b, b + NP2 * nlev);
  free(b);
  oracle_destroy(o);
}

void oracle_unit_vorticity(int nlev, const double *dvv, const double *D, const double *metdet, const double *v,
                           double *vort) {
  Oracle *o = unit_oracle(nlev, dvv, D, NULL, metdet, NULL, NULL, NULL);
  double *b = alloc0(2 * NP2 * nlev);
  vorticity_sphere(o, 0, nlev, v, v + NP2 * nlev, vort, b, b + NP2 * nlev);
  free(b);
  oracle_destroy(o);
}

void oracle_unit_laplace(int nlev, const double *dvv, const double *Dinv, const double *spheremp, const double *s,
                         double *lap) {
  Oracle *o = unit_oracle(nlev, dvv, NULL, Dinv, NULL, spheremp, NULL, NULL);
  double *b = alloc0(2 * NP2 * nlev);
  laplace_simple(o, 0, nlev, s, lap, b, b + NP2 * nlev);
  free(b);
  oracle_destroy(o);
}

void oracle_unit_vlaplace(int nlev, double nu_ratio, const double *dvv, const double *D, const double *Dinv,
                          const double *metdet, const double *spheremp, const double *mp, const double *metinv,
                          const double *v, double *lap) {
  Oracle *o = unit_oracle(nlev, dvv, D, Dinv, metdet, spheremp, mp, metinv);
  double *b = alloc0(6 * NP2 * nlev);
  vlaplace_sphere_wk_contra(o, 0, nlev, nu_ratio, v, v + NP2 * nlev, lap, lap + NP2 * nlev, b);
  free(b);
  oracle_destroy(o);
}

void oracle_unit_laplace_tensor(int nlev, const double *dvv, const double *Dinv, const double *spheremp,
                                const double *tensorvisc, const double *s, double *lap) {
  Oracle *o = unit_oracle(nlev, dvv, NULL, Dinv, NULL, spheremp, NULL, NULL);
  memcpy(o->tensorvisc, tensorvisc, sizeof(double) * 64);
  double *b = alloc0(4 * NP2 * nlev);
  laplace_tensor(o, 0, nlev, s, lap, b, b + NP2 * nlev, b + 2 * NP2 * nlev, b + 3 * NP2 * nlev);
  free(b);
  oracle_destroy(o);
}

void oracle_unit_vlaplace_cartesian(int nlev, const double *dvv, const double *Dinv, const double *spheremp,
                                    const double *tensorvisc, const double *vec_sph2cart, const double *v,
                                    double *lap) {
  Oracle *o = unit_oracle(nlev, dvv, NULL, Dinv, NULL, spheremp, NULL, NULL);
  memcpy(o->tensorvisc, tensorvisc, sizeof(double) * 64);
  memcpy(o->vec_sph2cart, vec_sph2cart, sizeof(double) * 96);
  double *b = alloc0(7 * NP2 * nlev);
  vlaplace_sphere_wk_cartesian(o, 0, nlev, v, v + NP2 * nlev, lap, lap + NP2 * nlev, b);
  free(b);
  oracle_destroy(o);
}

/* limiter on one (element, tracer): arrays in [pt][lev] layout */
void oracle_unit_limiter(int option, int nlev, const double *spheremp, const double *dpmass, double *qlim,
                         double *ptens) {
  limiter_shell(option, nlev, spheremp, dpmass, qlim, ptens);
}

/* PPM remap of nf columns (each nlev, contiguous) from src to tgt thickness */
void oracle_unit_ppm(int nlev, int remap_alg, const double *src, const double *tgt, int nf, double *vars) {
  double *dpo = alloc0(nlev + 8), *pio = alloc0(nlev + 8), *pin = alloc0(nlev + 8), *z2 = alloc0(nlev + 8);
  double *ppmdx = alloc0(10 * (nlev + 8)), *scr = alloc0(8 * (nlev + 8));
  int *kid = (int *)calloc(nlev + 8, sizeof(int));
  RemapGrid g = {dpo, pio, pin, z2, ppmdx, kid};
  ppm_grids(nlev, src, tgt, &g);
  for (int f = 0; f < nf; ++f) ppm_remap_column(nlev, remap_alg, &g, vars + (size_t)f * nlev, 1, scr);
  free(dpo); free(pio); free(pin); free(z2); free(ppmdx); free(scr); free(kid);
}

/* DSS of nf 3-D fields given in Fortran layout [ie][lev][np][np] (each field nelem*nlev*16) */
void oracle_unit_dss(Oracle *o, int nf, int nl, double *fields_f90, int apply_rspheremp) {
  if (nf <= 0) return;
  const size_t e3 = (size_t)NP2 * L;
  double *tmp = alloc0((size_t)nf * o->nelem * e3);
  double **f = (double **)malloc(sizeof(double *) * nf);
  size_t *es = (size_t *)malloc(sizeof(size_t) * nf);
  for (int i = 0; i < nf; ++i) {
    f[i] = tmp + (size_t)i * o->nelem * e3;
    es[i] = e3;
    for (int ie = 0; ie < o->nelem; ++ie) f2o3(fields_f90 + ((size_t)i * o->nelem + ie) * e3, f[i] + ie * e3, L);
  }
  dss(o, nf, f, es, nl, apply_rspheremp ? o->rspheremp : NULL);
  for (int i = 0; i < nf; ++i)
    for (int ie = 0; ie < o->nelem; ++ie) o2f3(f[i] + ie * e3, fields_f90 + ((size_t)i * o->nelem + ie) * e3, L);
  free(tmp); free(f); free(es);
}

int oracle_num_threads(void) { return nthreads(); }

void oracle_debug_stop(Oracle *o, int phase) { o->debug_stop = phase; }
