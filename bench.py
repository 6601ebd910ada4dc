"""Benchmark of the preqx step on one MI355X (FP64).

Workload: ne30 (5400 elements), np4, L72, 40 tracers, JW baroclinic jet with
perturbation, dt = 300 s, hyperviscosity subcycled 3x (nu = 1e15, nu_top =
2.5e5), remap every step, limiter 8, PPM remap alg 1, no forcing.  One step =
one prim_run_subcycle_c call through the C-ABI with all state resident in HBM.
The same steps are then run in the mixed-precision form from the same state
and its speed and per-field error against the FP64 run are reported, and the
cost of an adjoint step against a forward step is measured on the ne16 L72
Q4 differentiable-form workload (dry, rsplit = 1 and rsplit = 0).

Prints one JSON line: SYPD = dt / (365 * seconds per step), the roofline of the
dominant kernel (HIP-event durations on the library stream over the timed
steps; algorithmic bytes from the kernel's array traffic), the whole-step
fraction of HBM peak under the canonical bytes-per-step model, and a CPU
baseline (the oracle, OpenMP over elements, one step of the same workload).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu] [--no-mixed] [--no-adjoint]
                    [--ne 30] [--nlev 72] [--qsize 40] [--dt 300] [--nu 1e15] [--t-noise 0.0]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the phase whose launches the timed steps time with HIP events (the largest
# phase of this workload; the phase pass after the timed steps checks that)
DOMINANT_PHASE = "euler_advect"
METRIC = "SYPD @ne30 L72 Q40, 1 GPU: FP64 vs mixed (<=1e-6 rel); adjoint/fwd cost"


def f3_bytes(nelem, nlev):
    return nelem * 16 * nlev * 8


def step_model_bytes(nelem, nlev, q, h=3, r=1):
    """Canonical algorithmic bytes of one dynamics step (multiples of one FP64
    3-D field F3): prologue 3, RK5 109, HV H*(48 + 12/L), tracers 27.375Q + 57,
    remap (2Q + 8)/r, update_q 2Q/r."""
    b = 3 + 109 + h * (48 + 12.0 / nlev) + (27.375 * q + 57) + (2 * q + 8) / r + 2 * q / r
    return b * f3_bytes(nelem, nlev)


def kernel_bytes(name, nelem, nlev, q, launches):
    """Algorithmic bytes of one launch of a timed phase: every array the phase
    must read once plus every array it writes once (F3 = one FP64 3-D field,
    qlim = q/8 F3; element metric terms are < 0.1 F3 and left out).  Phases
    with several launches per step return the average over one step's
    launches."""
    F = f3_bytes(nelem, nlev)
    ql = q / 8.0
    if name == "euler_advect":
        # reads Qdp(Q), derived_dp, divdp_proj, vn0(2), divdp, fdss, qlim; writes qdp_pre(Q), fdss_pre.
        # Stage 2 also writes qlim; it reads stage 1's pre-exchange buffers (Q + 1) in place of the
        # assembled Qdp and divdp_proj and writes the assembled divdp_proj (its exchange is fused in).
        # Stage 3 also reads qbih(Q) and dpdiss_bih.
        s1 = (q + 6 + ql) + (q + 1)
        s2 = s1 + ql + 2
        s3 = s1 + q + 1
        return (s1 + s2 + s3) / 3.0 * F
    if name == "euler_dss":
        # two launches (stages 2 and 3): qdp_pre(Q) + fdss_pre in, Qdp(Q) + fdss out; stage 3 also
        # reads Qdp(n0)(Q) for the average
        s1 = 2 * q + 2
        return (2 * s1 + q) / 2.0 * F
    if name == "euler_qmm_lap":
        # Qdp(Q), derived_dp, divdp_proj, dpdiss_ave in; qbih_pre(Q), qlim out
        return (2 * q + 3 + ql) * F
    if name == "euler_bih_post":
        return 2 * q * F
    if name == "euler_divdp_qmm":
        return (q + 3 + 2 + ql) * F
    if name == "caar":
        return (10 + 3 + 1 + 4) * F  # reads n0(4), nm1(4), vn0(2), omega; writes 4, vn0(2), omega, phi
    if name == "caar_dss":
        return 8 * F
    if name == "remap":
        return ((3 + q) * 2 + 1 + q) * F  # v,T,Qdp in/out, dp3d, Q out
    return None


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "preqx_step_ne30_L72_Q40_hbm_traffic.json")


def measured_traffic(phase, args):
    """HBM bytes per launch of a timed phase from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (measured on the default ne30 L72 Q40
    workload only; None for other sizes or phases)."""
    if (args.ne, args.nlev, args.qsize) != (30, 72, 40) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        return json.load(f)["traffic_bytes_per_launch"].get(phase)


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_problem(args):
    from hommexx_amd.problem import build_problem
    nu = args.nu
    return build_problem(ne=args.ne, nlev=args.nlev, qsize=args.qsize, seed=1, dt=args.dt, nu=nu, nu_p=nu,
                         nu_q=nu, nu_s=nu, nu_div=nu, nu_top=2.5e5, hypervis_subcycle=3, rsplit=1,
                         t_noise=args.t_noise)


def run_gpu(args, prob, form="fp64"):
    import numpy as np
    from hommexx_amd.hip import HipSession

    progress("problem: nelem %d nlev %d qsize %d, form %s" % (prob.nelem, prob.nlev, prob.qsize, form))
    sph = prob.metric["spheremp"]
    with HipSession(prob, abort_on_error=True, form=form) as h:
        progress("state resident on the GPU")
        h.run(args.warmup)
        h.synchronize()
        progress("%d warmup steps done" % args.warmup)
        # the timed steps carry HIP events around the launches of one phase
        # only (the dominant one); every phase is timed in a second pass
        h.profile_reset()
        h.profile(True, only=DOMINANT_PHASE)
        t0 = time.perf_counter()
        h.run(args.steps)
        h.synchronize()
        t1 = time.perf_counter()
        progress("%d timed steps done: %.3f ms/step" % (args.steps, (t1 - t0) / args.steps * 1e3))
        h.profile(False)
        prof_timed = h.profile_report()
        h.profile_reset()
        h.profile(True)
        h.run(args.steps)
        h.synchronize()
        h.profile(False)
        prof = h.profile_report()
        progress("%d more steps with every phase timed" % args.steps)
        st = h.state()
        nstep, nm1, n0, np1 = h.tl
    sec = (t1 - t0) / args.steps
    prof = {"timed": prof_timed, "phases": prof}
    # conservation check over warmup + timed steps
    dry0 = float((sph * prob.state["ps_v"][:, 1]).sum())
    dry = float((sph * st["ps_v"][:, n0]).sum())
    q0 = np.einsum("epq,eclpq->c", sph, prob.state["Qdp"][:, 0, :args.qsize])
    q1 = np.einsum("epq,eclpq->c", sph, st["Qdp"][:, nstep % 2, :args.qsize])
    return sec, prof, abs(dry / dry0 - 1.0), float(np.max(np.abs(q1 / q0 - 1.0))), st


def run_cpu(prob, min_seconds=10.0, max_steps=20):
    """The oracle on the host cores: steps of the same workload until about
    ``min_seconds`` of CPU work (at least one step)."""
    import oracle
    oracle.build()
    s = oracle.OracleSession(prob)
    t0 = time.perf_counter()
    s.run(1)
    n = 1
    while time.perf_counter() - t0 < min_seconds and n < max_steps:
        s.run(1)
        n += 1
    t1 = time.perf_counter()
    progress("cpu baseline: %d steps in %.2f s" % (n, t1 - t0))
    threads = oracle.lib().oracle_num_threads()
    s.close()
    return (t1 - t0) / n, n, threads


def run_adjoint(rsplit, steps=10, traj_steps=48, every=8):
    """Cost of the differentiable form on the ne16 L72 Q4 workload (dry,
    dt = 300 s, ``rsplit`` 0 or 1): one adjoint step (hommexx_hip_adjoint_step,
    which recomputes the step's stage states and runs the reverse sweep)
    against one forward step (prim_run_subcycle_c), both timed over
    ``steps`` calls after a warm-up; and a checkpointed 48-step trajectory,
    forward and backward through torch.autograd."""
    import torch
    from hommexx_amd.adjoint import (DiffSession, energy_mismatch, initial_state, region_weight,
                                     synthetic_target, trajectory)
    from hommexx_amd.problem import build_problem
    nu = 7e15
    prob = build_problem(ne=16, nlev=72, qsize=4, seed=32, rsplit=rsplit, nu=nu, nu_p=nu, nu_q=nu, nu_s=nu,
                         nu_div=nu)
    progress("adjoint: ne16 L72 Q4 rsplit=%d, nelem %d" % (rsplit, prob.nelem))
    with DiffSession(prob) as s:
        s.step(2)
        s.h.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.h.synchronize()
        fwd = (time.perf_counter() - t0) / steps
        st = s.get_state()
        gen = torch.Generator(device="cuda").manual_seed(3)
        lam = [torch.randn(a.shape, generator=gen, dtype=torch.float64, device="cuda") for a in st]
        s.adjoint_step(st, lam)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.adjoint_step(st, lam)
        torch.cuda.synchronize()
        adj = (time.perf_counter() - t0) / steps
        # the dry dynamics alone (no tracer step): the map the adjoint differentiates
        y = tuple(a.clone() for a in st)
        s.dynamics_step(y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.dynamics_step(y)
        torch.cuda.synchronize()
        dyn = (time.perf_counter() - t0) / steps
        progress("adjoint: forward %.3f ms/step (dynamics only %.3f), adjoint %.3f ms/step"
                 % (fwd * 1e3, dyn * 1e3, adj * 1e3))
        tgt, w = synthetic_target(prob, 6), region_weight(prob)
        # warm-up: a 2-step trajectory and its gradient (first-use costs of
        # the loss's torch kernels and of the checkpoint allocations)
        x = [a.clone().requires_grad_() for a in initial_state(prob)]
        energy_mismatch(prob, *trajectory(s, *x, 2, 1), tgt, w).backward()
        x = [a.clone().requires_grad_() for a in initial_state(prob)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = energy_mismatch(prob, *trajectory(s, *x, traj_steps, every), tgt, w)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    progress("adjoint: trajectory forward %.3f s, backward %.3f s" % (t1 - t0, t2 - t1))
    vert = "rsplit=0 (Eulerian vertical)" if rsplit == 0 else "rsplit=1 (PPM remap of the states every step)"
    return {"config": {"workload": "preqx step ne16 np4 L72 Q4, dt=300s, %s, dry, hv subcycle 3, nu=7e15" % vert,
                       "nelem": prob.nelem},
            "forward_step_ms": round(fwd * 1e3, 3), "adjoint_step_ms": round(adj * 1e3, 3),
            "adjoint_over_forward": round(adj / fwd, 3),
            "dynamics_step_ms": round(dyn * 1e3, 3), "adjoint_over_dynamics": round(adj / dyn, 3),
            "trajectory": {"steps": traj_steps, "checkpoint_every": every,
                           "forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4),
                           "backward_over_forward": round((t2 - t1) / (t1 - t0), 3)}}


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters: each launch reads and writes ``nbytes``."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ne", type=int, default=30)
    ap.add_argument("--nlev", type=int, default=72)
    ap.add_argument("--qsize", type=int, default=40)
    ap.add_argument("--dt", type=float, default=300.0)
    ap.add_argument("--nu", type=float, default=1e15, help="hyperviscosity (all of nu, nu_p, nu_q, nu_s, nu_div)")
    ap.add_argument("--t-noise", type=float, default=0.0,
                    help="seeded white-noise temperature perturbation (K) on the initial state")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision form")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-cost leg")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint-cost leg (rsplit 1 and 0); prints its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the preqx step here is built and measured on a single MI355X; "
              "multi-GPU runs are not part of this build (the step would shard by elements, with one halo "
              "exchange per DSS)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.adjoint_only:
        a1 = run_adjoint(1)
        a1["rsplit0"] = run_adjoint(0)
        print(json.dumps({"adjoint": a1}))
        return
    progress("building the ne%d L%d Q%d problem on the host" % (args.ne, args.nlev, args.qsize))
    prob = make_problem(args)
    sec, prof, dry_drift, q_drift, st64 = run_gpu(args, prob, "fp64")
    nelem, nlev, q = prob.nelem, prob.nlev, args.qsize
    mixed = None
    if not args.no_mixed:
        import numpy as np
        msec, _, mdry, mq, stm = run_gpu(args, prob, "mixed")
        err = {k: float(np.linalg.norm((stm[k] - st64[k]).ravel()) / np.linalg.norm(st64[k].ravel()))
               for k in ("v", "T", "dp3d", "ps_v", "Qdp", "Q")}
        mixed = {"ms_per_step": round(msec * 1e3, 3), "value": round(args.dt / (365.0 * msec), 3),
                 "speedup_vs_fp64": round(sec / msec, 4), "max_rel_l2_vs_fp64": max(err.values()),
                 "rel_l2_vs_fp64": err, "steps_compared": args.warmup + 2 * args.steps,
                 "dry_mass_rel_drift": mdry, "tracer_mass_rel_drift_max": mq}
    adjoint = None
    if not args.no_adjoint:
        a1, a0 = run_adjoint(1), run_adjoint(0)
        adjoint = dict(a1)
        adjoint["rsplit0"] = a0
    sypd = args.dt / (365.0 * sec)
    # dominant kernel by total time (phase pass); its average launch time from
    # the timed steps when it is the phase they timed
    kern = {k: v for k, v in prof["phases"].items() if kernel_bytes(k, nelem, nlev, q, 1) is not None}
    dom = max(kern, key=lambda k: kern[k][1]) if kern else None
    roof = None
    if dom:
        timed = dom in prof["timed"]
        cnt, tot = prof["timed"][dom] if timed else kern[dom]
        avg_s = tot / cnt / 1e3
        byts = kernel_bytes(dom, nelem, nlev, q, cnt)
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, args), "bytes_per_launch": byts,
                "avg_launch_ms": round(avg_s * 1e3, 4), "launches": cnt,
                "measured_in": "timed steps" if timed else "phase pass"}
    model = step_model_bytes(nelem, nlev, q)
    step_roof = {"model_bytes_per_step": model, "achieved": round(model / sec / 1e9, 1),
                 "frac": round(model / sec / 1e9 / HBM_PEAK_GBS, 4)}
    phases = {k: {"launches": v[0], "ms_per_step": round(v[1] / args.steps, 4)}
              for k, v in sorted(prof["phases"].items())}
    cpu = None
    if not args.no_cpu:
        csec, nsteps, threads = run_cpu(prob)
        cpu = {"value": round(args.dt / (365.0 * csec), 6), "unit": "SYPD", "cores": threads, "kind": "port",
               "sample": "%d prim_run_subcycle_c steps of the same ne%d L%d Q%d workload on the oracle "
                         "(C, OpenMP over elements): %.3f s/step" % (nsteps, args.ne, nlev, q, csec)}
    out = {
        "metric": METRIC, "value": round(sypd, 3), "unit": "SYPD", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "preqx step ne%d np4 L%d Q%d, dt=%gs, nu=%g, hv subcycle 3, remap every step, "
                               "limiter 8, JW baroclinic%s" % (args.ne, nlev, q, args.dt, args.nu,
                                                               " + %g K T noise" % args.t_noise if args.t_noise else ""),
                   "nelem": nelem, "nlev": nlev, "qsize": q, "form": "fp64"},
        "roofline": roof, "step_roofline": step_roof, "cpu_baseline": cpu,
        "conservation": {"dry_mass_rel_drift": dry_drift, "tracer_mass_rel_drift_max": q_drift},
        "mixed": mixed,
        "adjoint": adjoint,
        "phases": phases,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
