/*
 * hommexx_amd.h -- C-ABI of libhommexx_amd.so, the gfx950 HIP implementation
 * of the HOMMEXX preqx time step.
 *
 * The first group is the reference's own iso_c_binding interface, with the
 * same names, argument order and argument meaning (paths relative to the
 * reference root).  Fortran passes scalars by reference and arrays as
 * references to c_ptr values pointing at the flat elem_* arrays, so C callers
 * pass the address of a pointer.  Array layouts are the Fortran ones:
 *   v      (np,np,2,nlev,timelevels=3,nelemd)    T, dp3d (np,np,nlev,3,nelemd)
 *   Qdp    (np,np,nlev,qsize_d,2,nelemd)          Q (np,np,nlev,qsize_d,nelemd)
 *   ps_v   (np,np,3,nelemd)                       omega_p (np,np,nlev,nelemd)
 *   D, Dinv, metinv, tensorvisc (np,np,2,2,nelemd); vec_sph2cart (np,np,3,2,nelemd)
 *   fcor, mp, spheremp, rspheremp, metdet, phis (np,np,nelemd); dvv (np,np)
 * Time-level indices enter 1-based (init_time_level_c) and leave
 * prim_run_subcycle_c 0-based, as in the reference.  Errors the reference
 * turns into MPI_Abort abort the process unless
 * hommexx_hip_set_abort_on_error(0) was called.
 */
#ifndef HOMMEXX_AMD_H
#define HOMMEXX_AMD_H

#ifdef __cplusplus
#include <cstdbool>
extern "C" {
#define HX_REF(T) T &
#else
#include <stdbool.h>
#define HX_REF(T) T *
#endif

/* ---------------------------------------------------------------------------
 * Reference C-ABI (18 symbols)
 * ------------------------------------------------------------------------- */

/* src/share/cxx/Hommexx_Session.cpp:35 -- create the device stream */
void initialize_hommexx_session(void);
/* src/share/cxx/Hommexx_Session.cpp:75 -- release every device buffer */
void finalize_hommexx_session(void);

/* src/share/cxx/mpi/mpi_cxx_f90_interface.cpp:20 (one rank: recorded only) */
void reset_cxx_comm(HX_REF(const int) f_comm);
/* src/share/cxx/mpi/mpi_cxx_f90_interface.cpp:27 */
void init_connectivity(HX_REF(const int) num_local_elems);
/* src/share/cxx/mpi/mpi_cxx_f90_interface.cpp:34-55: 1-based ids; positions
 * in the Fortran order W,E,S,N,SW,SE,NW,NE */
void add_connection(HX_REF(const int) first_lid, HX_REF(const int) first_gid, HX_REF(const int) first_pos,
                    HX_REF(const int) first_pid, HX_REF(const int) second_lid, HX_REF(const int) second_gid,
                    HX_REF(const int) second_pos, HX_REF(const int) second_pid);
/* src/share/cxx/mpi/mpi_cxx_f90_interface.cpp:57 -- builds the DSS gather tables */
void finalize_connectivity(void);

/* src/share/cxx/cxx_f90_interface.cpp:207 */
void init_derivative_c(HX_REF(const double *const) dvv);
/* src/share/cxx/cxx_f90_interface.cpp:33-117 (same option checks) */
void init_simulation_params_c(HX_REF(const int) remap_alg, HX_REF(const int) limiter_option, HX_REF(const int) rsplit,
                              HX_REF(const int) qsplit, HX_REF(const int) time_step_type,
                              HX_REF(const int) energy_fixer, HX_REF(const int) qsize,
                              HX_REF(const int) state_frequency, HX_REF(const double) nu, HX_REF(const double) nu_p,
                              HX_REF(const double) nu_q, HX_REF(const double) nu_s, HX_REF(const double) nu_div,
                              HX_REF(const double) nu_top, HX_REF(const int) hypervis_order,
                              HX_REF(const int) hypervis_subcycle, HX_REF(const double) hypervis_scaling,
                              HX_REF(const int) ftype, HX_REF(const bool) prescribed_wind,
                              HX_REF(const bool) moisture, HX_REF(const bool) disable_diagnostics,
                              HX_REF(const bool) use_cpstar, HX_REF(const bool) use_semi_lagrangian_transport);
/* src/share/cxx/cxx_f90_interface.cpp:224-233 */
void init_elements_2d_c(HX_REF(const int) num_elems, HX_REF(const double *const) D, HX_REF(const double *const) Dinv,
                        HX_REF(const double *const) fcor, HX_REF(const double *const) mp,
                        HX_REF(const double *const) spheremp, HX_REF(const double *const) rspheremp,
                        HX_REF(const double *const) metdet, HX_REF(const double *const) metinv,
                        HX_REF(const double *const) phis, HX_REF(const double *const) tensorvisc,
                        HX_REF(const double *const) vec_sph2cart, HX_REF(const bool) consthv);
/* src/share/cxx/cxx_f90_interface.cpp:235-251 */
void init_elements_states_c(HX_REF(const double *const) v, HX_REF(const double *const) T,
                            HX_REF(const double *const) dp3d, HX_REF(const double *const) qdp,
                            HX_REF(const double *const) ps_v);
/* src/share/cxx/cxx_f90_interface.cpp:253-262, Diagnostics.cpp:17-35: the
 * Fortran arrays q (np,np,nlev,qsize_d), qvar and qmass (np,np,qsize_d,4),
 * q1mass (np,np,qsize_d), iener, kener, pener (np,np,4), iener_wet (np,np),
 * per element; prim_run_subcycle_c fills them on the reference's diagnostic
 * steps unless disable_diagnostics (prim_driver.cpp:49-63,84-88,140-144) */
void init_diagnostics_c(HX_REF(double *) q, HX_REF(double *) qvar, HX_REF(double *) qmass, HX_REF(double *) q1mass,
                        HX_REF(double *) iener, HX_REF(double *) iener_wet, HX_REF(double *) kener,
                        HX_REF(double *) pener);
/* src/share/cxx/cxx_f90_interface.cpp:119-124 (hyam, hybm unused as in the reference) */
void init_hvcoord_c(HX_REF(const double) ps0, HX_REF(const double *const) hybrid_am,
                    HX_REF(const double *const) hybrid_ai, HX_REF(const double *const) hybrid_bm,
                    HX_REF(const double *const) hybrid_bi);
/* src/share/cxx/cxx_f90_interface.cpp:264-280 -- allocates exchange/scratch buffers */
void init_boundary_exchanges_c(void);
/* src/share/cxx/cxx_f90_interface.cpp:213-222 (1-based in) */
void init_time_level_c(HX_REF(const int) nm1, HX_REF(const int) n0, HX_REF(const int) np1, HX_REF(const int) nstep,
                       HX_REF(const int) nstep0);
/* src/share/cxx/prim_driver.cpp:31-156 -- one remap interval; outputs 0-based */
void prim_run_subcycle_c(HX_REF(const double) dt, HX_REF(int) nstep, HX_REF(int) nm1, HX_REF(int) n0,
                         HX_REF(int) np1, HX_REF(const int) last_time_step);
/* src/share/cxx/cxx_f90_interface.cpp:126-154 */
void cxx_push_results_to_f90(HX_REF(double *) v, HX_REF(double *) T, HX_REF(double *) dp3d, HX_REF(double *) Qdp,
                             HX_REF(double *) Q, HX_REF(double *) ps_v, HX_REF(double *) omega_p);
/* src/share/cxx/cxx_f90_interface.cpp:157 (FM (np,np,2,nlev,nelemd), FT (np,np,nlev,nelemd),
 * FQ (np,np,nlev,qsize_d,nelemd)) */
void cxx_push_forcing_to_f90(double *FM, double *FT, double *FQ);
/* src/share/cxx/cxx_f90_interface.cpp:180 */
void f90_push_forcing_to_cxx(double *FM, double *FT, double *FQ, double *Qdp);

/* ---------------------------------------------------------------------------
 * Extensions (no counterpart in the reference; for the Python harness,
 * tests and benchmarks)
 * ------------------------------------------------------------------------- */

/* nlev and qsize_d are compile-time constants in the reference (PLEV,
 * QSIZE_D; cmake -DPREQX_PLEV, -DQSIZE_D, defaults 72 and 4).  The library
 * takes its defaults from the same build variables (make PREQX_PLEV=...
 * QSIZE_D=...), so an unchanged Fortran driver needs no call here; a harness
 * may override them before init_elements_2d_c (qsize_d <= 0: qsize).
 * Returns 0, 1 if elements already exist, 2 if nlev is outside 3..256. */
int hommexx_hip_configure(int nlev, int qsize_d);
/* the nlev and qsize_d the next init_elements_2d_c will use */
void hommexx_hip_dimensions(int *nlev, int *qsize_d);
/* 1 (default): abort like the reference; 0: record the error and return */
void hommexx_hip_set_abort_on_error(int on);
/* 0 (default): the FP64 step.  1: the mixed-precision form -- the sphere
 * operators of hyperviscosity, of the tracer flux divergence and of the
 * tracer biharmonic run in FP32 with compensated contractions, and the
 * tracer values handed from one kernel to the next exchange (pre-exchange
 * buffers) are stored in FP32; the prognostic state, column scans, remap,
 * DSS accumulation, limiters and RK combinations stay FP64.  Takes effect at
 * the next prim_run_subcycle_c.  Returns 1 for an unknown form. */
int hommexx_hip_set_form(int form);
/* debugging aid: prim_run_subcycle_c returns after phase n (1: first RK stage,
 * 2: RK stages, 3: hyperviscosity, 4: first tracer stage, 5: tracers), 0: off */
void hommexx_hip_debug_stop(int phase);
int hommexx_hip_last_error(void);
const char *hommexx_hip_error_message(void);
void hommexx_hip_clear_error(void);
void hommexx_hip_synchronize(void);
/* the hipStream_t every kernel is launched on */
void *hommexx_hip_stream(void);
/* HIP-event timing of each step phase on the library stream */
void hommexx_hip_profile(int enable);
/* time only the phase of this name (NULL or "": every phase): fewer events in
   the timed steps */
void hommexx_hip_profile_only(const char *name);
void hommexx_hip_profile_reset(void);
int hommexx_hip_profile_get(const char *name, int *count, double *total_ms);
int hommexx_hip_profile_names(char *buf, int len);
void hommexx_hip_time_levels(int *nm1, int *n0, int *np1, int *nstep, int *n0_qdp, int *np1_qdp);

/* Differentiable form (the reference has no adjoint; these serve the
 * torch.autograd harness hommexx_amd/adjoint.py).  Device pointers in the
 * library's layouts: v [ie][lev][2][pt], T [ie][lev][pt], ps [ie][pt]
 * (pt = igp*4 + jgp).  get/set copy the prognostic state of time level n0
 * (set also recomputes dp3d(n0) from ps).  hommexx_hip_adjoint_step takes
 * the input state of one prim_run_subcycle_c step of length dt and, in
 * place, maps the adjoint of the step's output (v, T, ps) to the adjoint of
 * its input (all qsplit * max(rsplit, 1) dynamics steps and the remap); it
 * needs a dry configuration, constant-coefficient hyperviscosity, no forcing
 * and the FP64 form, and leaves the session's model state unchanged.
 * Return 0 on success. */
int hommexx_hip_get_state(double *v, double *T, double *ps);
/* one field of the current state into device memory: which 0 = v, 1 = T,
 * 2 = dp3d, 3 = ps_v (time level n0), 4 = Qdp of tracer q (the tracer level
 * the next step starts from), 5 = Q of tracer q; layouts as above, tracer
 * fields [ie][lev][pt].  For comparisons on the device (no host copies of
 * large states).  Returns 0 on success. */
int hommexx_hip_get_field(int which, int q, double *dst);
int hommexx_hip_set_state(const double *v, const double *T, const double *ps);
int hommexx_hip_adjoint_step(double dt, const double *v, const double *T, const double *ps, double *lam_v,
                             double *lam_T, double *lam_ps);
/* the adjoint of n consecutive steps of length dt (one checkpointed segment)
 * from the state (v, T, ps) before the first; lam_* as for
 * hommexx_hip_adjoint_step.  The forward pass keeps the stage states of every
 * dynamics step of the segment (40 3-D fields each at hypervis_subcycle 3),
 * so nothing is recomputed in the reverse sweep.  Same conditions; the
 * session's model state is not changed.  Returns 0 on success. */
int hommexx_hip_adjoint_segment(double dt, int n, const double *v, const double *T, const double *ps, double *lam_v,
                                double *lam_T, double *lam_ps);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving or training loops).
 * hommexx_hip_adjoint_step with every buffer taken from one fixed device
 * array of 16 Mi doubles (128 MiB) that the library holds on each GPU until
 * the process ends; it calls no hipMalloc or hipFree.  Configurations that
 * need more (about ne4 at 72 levels is the limit) fail with error 71 (return
 * 3); init_boundary_exchanges_c must have run.  All calls on one GPU share
 * its array: calls that overlap on one GPU race.  Same results as
 * hommexx_hip_adjoint_step. */
int hommexx_hip_adjoint_step_noalloc(double dt, const double *v, const double *T, const double *ps, double *lam_v,
                                     double *lam_T, double *lam_ps);
/* the map hommexx_hip_adjoint_step differentiates, forward: (v, T, ps) is
 * advanced in place by one prim_run_subcycle_c step of length dt without its
 * tracer step (for a dry run the (v, T, ps) that step produces); used to
 * replay checkpointed segments.  Same conditions; the session's model state
 * is not changed.  Returns 0 on success. */
int hommexx_hip_dynamics_step(double dt, double *v, double *T, double *ps);

#ifdef __cplusplus
}
#endif
#undef HX_REF
#endif /* HOMMEXX_AMD_H */
