"""Whole-step parity at the benchmark size: ne30 (5400 elements), 72 levels,
40 tracers, dt 300 s, hyperviscosity nu 1e15 subcycled 3 times, remap every
step, limiter 8 -- the workload of the reference's ne30 benchmark namelist
(test/reg_test/benchmarks/v1/xx-ne30-v1.nl) with rsplit 1.

One oracle session and one HIP session advance side by side through the
tests of this module, in order: 1 step (bar 1e-13 relative L2), 10 steps
(bar 1e-10), then on to 100 steps in chunks, so that no test runs for more
than about a minute.  At the end, the dry and tracer mass drifts of the HIP
run are compared with the oracle's over the same 100 steps.
"""
import numpy as np
import pytest

import oracle
from hommexx_amd.hip import HipSession
from hommexx_amd.problem import build_problem
from parity import compare, fmt

pytestmark = pytest.mark.gpu

NU = 1e15


@pytest.fixture(scope="module")
def pair():
    prob = build_problem(ne=30, nlev=72, qsize=40, seed=1, dt=300.0, nu=NU, nu_p=NU, nu_q=NU, nu_s=NU, nu_div=NU,
                         nu_top=2.5e5, hypervis_subcycle=3, rsplit=1)
    o = oracle.OracleSession(prob)
    h = HipSession(prob)
    state = {"steps": 0}
    yield prob, o, h, state
    h.close()
    o.close()


def advance(pair, to):
    prob, o, h, state = pair
    n = to - state["steps"]
    assert n > 0
    o.run(n)
    h.run(n)
    state["steps"] = to
    return h.state(), o.state()


def masses(prob, st, tl):
    sph = prob.metric["spheremp"]
    nstep, nm1, n0, np1 = tl
    dry = float((sph * st["ps_v"][:, n0]).sum())
    q = np.einsum("epq,eclpq->c", sph, st["Qdp"][:, nstep % 2, :prob.qsize])
    return dry, q


@pytest.mark.parametrize("to,tol", [(1, 1e-13), (10, 1e-10), (40, 1e-8), (70, 1e-8), (100, 1e-8)])
def test_ne30_l72_q40(pair, to, tol):
    out, ref = advance(pair, to)
    rep = compare(out, ref)
    print("ne30 L72 Q40 after %d steps:" % to, fmt(rep), flush=True)
    for k, (e, _) in rep.items():
        assert e <= tol, (to, k, e)
    if to == 100:
        prob, o, h, _ = pair
        dry0, q0 = masses(prob, prob.state, (0, 0, 1, 2))
        dh, qh = masses(prob, out, h.tl)
        do, qo = masses(prob, ref, o.tl)
        print("dry mass drift: HIP %.2e oracle %.2e" % (dh / dry0 - 1, do / dry0 - 1))
        print("tracer mass drift (max over tracers): HIP %.2e oracle %.2e"
              % (np.max(np.abs(qh / q0 - 1)), np.max(np.abs(qo / q0 - 1))))
        assert abs(dh / dry0 - 1) <= 1e-13
        assert np.max(np.abs(qh / q0 - 1)) <= 1e-12
        assert np.max(np.abs(qh - qo) / q0) <= 1e-13
