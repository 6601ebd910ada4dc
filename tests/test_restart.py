"""Restart through the C-ABI: a run that pushes its state back to the
Fortran arrays (cxx_push_results_to_f90), starts a new session from them
(init_elements_states_c, init_time_level_c with the 1-based levels the
Fortran driver adds back, prim_main.F90:305-308) and continues, matches the
uninterrupted run bit for bit.  This is what a Fortran restart
(restart_io_mod.F90, prim_main.F90:355-357) relies on: the step carries no
state beyond the prognostic arrays and the time levels.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from hommexx_amd.problem import build_problem

FIELDS = ("v", "T", "dp3d", "Qdp", "ps_v", "Q")


def restarted(prob, state, tl):
    """``prob`` restarted from pushed Fortran arrays ``state`` and the 0-based
    time levels (nstep, nm1, n0, np1) returned by prim_run_subcycle_c"""
    nstep, nm1, n0, np1 = tl
    s = dict(prob.state)
    for k in ("v", "T", "dp3d", "Qdp", "ps_v"):
        s[k] = state[k].copy()
    return dataclasses.replace(prob, state=s, time_level=(nm1 + 1, n0 + 1, np1 + 1, nstep, prob.time_level[4]))


def cases():
    return [dict(), dict(rsplit=3, qsplit=2), dict(moisture=True, limiter_option=9)]


@pytest.mark.parametrize("over", cases())
def test_oracle_restart_is_exact(over):
    prob = build_problem(ne=2, nlev=26, qsize=3, seed=6, **over)
    a = oracle.OracleSession(prob)
    tl = a.run(2)
    mid = a.state()
    a.run(2)
    ref = a.state()
    a.close()
    b = oracle.OracleSession(restarted(prob, mid, tl))
    b.run(2)
    out = b.state()
    b.close()
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("over", cases())
def test_restart_is_exact(over):
    from hommexx_amd.hip import HipSession
    prob = build_problem(ne=4, nlev=26, qsize=4, seed=6, **over)
    print("case: ne4 L26 Q4 restart after 2 steps, %r" % (over,), flush=True)
    with HipSession(prob) as h:
        tl = h.run(2)
        mid = h.state()
        tl2 = h.run(2)
        ref = h.state()
    with HipSession(restarted(prob, mid, tl)) as h:
        assert h.run(2) == tl2
        out = h.state()
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
