"""Run preqx steps of the benchmark workload through the C-ABI, in one form.

    python -m hommexx_amd.run_steps [--form fp64|mixed] [--steps 10] [--warmup 2]
                                    [--ne 30] [--nlev 72] [--qsize 40]

Builds the jet initial state of the ne30 np4 L72 Q40 workload (dt 300 s, all
hyperviscosity coefficients 1e15, three subcycles, remap every step), hands it
to the library, runs the warm-up steps and then the timed steps, and prints
the wall time per step.  A plain driver for profilers and for users who want
the step alone, without the comparisons of bench.py.
"""
import argparse
import time

from .hip import HipSession
from .problem import build_problem


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--form", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ne", type=int, default=30)
    ap.add_argument("--nlev", type=int, default=72)
    ap.add_argument("--qsize", type=int, default=40)
    ap.add_argument("--dt", type=float, default=300.0)
    ap.add_argument("--nu", type=float, default=1e15)
    args = ap.parse_args()
    nu = args.nu
    prob = build_problem(ne=args.ne, nlev=args.nlev, qsize=args.qsize, seed=1, dt=args.dt, nu=nu, nu_p=nu, nu_q=nu,
                         nu_s=nu, nu_div=nu, nu_top=2.5e5, hypervis_subcycle=3, rsplit=1)
    with HipSession(prob, abort_on_error=True, form=args.form) as h:
        h.run(args.warmup)
        h.synchronize()
        t0 = time.perf_counter()
        h.run(args.steps)
        h.synchronize()
        t1 = time.perf_counter()
    print("%s ne%d L%d Q%d: %d steps, %.3f ms/step" % (args.form, args.ne, args.nlev, args.qsize, args.steps,
                                                      (t1 - t0) / args.steps * 1e3), flush=True)


if __name__ == "__main__":
    main()
