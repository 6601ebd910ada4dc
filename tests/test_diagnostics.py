"""Diagnostics (Diagnostics.cpp:37-185, scheduled by prim_driver.cpp:42-63,
84-88, 140-144): the per-column energy and tracer sums that
prim_run_subcycle_c writes into the Fortran arrays of init_diagnostics_c.

CPU tests pin the oracle's diagnostics:
* PEner has a closed form: phis times the column's total thickness;
* the slot schedule: 3 and 2 before the advance, 0 after the (absent)
  forcing, 1 after the step;
* the Qmass sums carry tracer-mass conservation from slot 3 to slot 1;
* the state-frequency / last-step schedule.

Each check fails on an oracle build with a planted mistake (13: PEner
doubled, 14: the first scalar call writes slot 2).  GPU tests compare the HIP
library's arrays with the oracle's.
"""
import numpy as np
import pytest

import oracle
from hommexx_amd.problem import build_problem

SLOTS_SET = {"qvar": (0, 1, 3), "qmass": (0, 1, 3), "iener": (0, 1, 2), "kener": (0, 1, 2), "pener": (0, 1, 2)}


def diag_problem(**kw):
    kw.setdefault("state_frequency", 1)
    return build_problem(ne=2, nlev=26, qsize=3, seed=4, disable_diagnostics=False, **kw)


def oracle_diag(prob, nsteps=1, L=None, last_time_step=1 << 30):
    o = oracle.OracleSession(prob, L=L)
    o.run(nsteps, last_time_step=last_time_step)
    d = o.diagnostics()
    o.close()
    return d


def pener_closed_form_error(L=None):
    """max relative error of PEner slot 2 (before the first step) against
    phis * (ps0 (hyai(L) - hyai(0)) + ps (hybi(L) - hybi(0)))"""
    prob = diag_problem()
    d = oracle_diag(prob, 1, L)
    vc = prob.vcoord
    ps = prob.state["ps_v"][:, prob.time_level[1] - 1]
    thick = vc.ps0 * (vc.hyai[-1] - vc.hyai[0]) + ps * (vc.hybi[-1] - vc.hybi[0])
    want = prob.state["phis"] * thick
    return np.abs(d["pener"][:, 2] - want).max() / np.abs(want).max()


def test_oracle_pener_closed_form():
    assert pener_closed_form_error() < 1e-13


def test_pener_check_detects_planted_factor(oracle_lib):
    assert not pener_closed_form_error(oracle.plant(13)) < 1e-13


def slot_errors(L=None):
    """(|qmass slot 3 - slot 0|, |iener slot 2 - slot 0|, relative global
    tracer-mass change from slot 3 to slot 1, |q1mass - qmass slot 1|)"""
    prob = diag_problem()
    d = oracle_diag(prob, 1, L)
    sph = prob.metric["spheremp"][:, None]
    m3 = (d["qmass"][:, 3, :prob.qsize] * sph).sum(axis=(0, 2, 3))
    m1 = (d["qmass"][:, 1, :prob.qsize] * sph).sum(axis=(0, 2, 3))
    return (np.abs(d["qmass"][:, 3] - d["qmass"][:, 0]).max(), np.abs(d["iener"][:, 2] - d["iener"][:, 0]).max(),
            np.abs(m1 - m3).max() / max(np.abs(m3).max(), 1e-300), np.abs(d["q1mass"] - d["qmass"][:, 1]).max())


def test_oracle_slots_and_tracer_mass():
    q30, e20, dm, q1 = slot_errors()
    assert q30 == 0 and e20 == 0 and q1 == 0
    assert dm < 1e-13


def test_slot_check_detects_planted_slot(oracle_lib):
    q30, _, dm, _ = slot_errors(oracle.plant(14))
    assert not (q30 == 0 and dm < 1e-13)


def test_oracle_qvar_and_q_copy():
    """slot 1 (after the step): Qvar = sum_k Qdp Q at the tracer level whose
    level sum is Qmass, and the q array is the model's Q"""
    prob = diag_problem()
    o = oracle.OracleSession(prob)
    o.run(1)
    d, s = o.diagnostics(), o.state()
    o.close()
    np.testing.assert_array_equal(d["q"], s["Q"])
    qd = prob.qsize
    lev = [tl for tl in (0, 1) if np.allclose(s["Qdp"][:, tl, :qd].sum(axis=2), d["qmass"][:, 1, :qd], rtol=1e-14)]
    assert len(lev) == 1
    want = (s["Qdp"][:, lev[0], :qd] * s["Q"][:, :qd]).sum(axis=2)
    assert np.abs(d["qvar"][:, 1, :qd] - want).max() <= 1e-13 * np.abs(want).max()


def test_oracle_schedule():
    # state_frequency 3: the third step computes, the first two do not
    prob = diag_problem(state_frequency=3)
    assert all(np.all(v == 0) for k, v in oracle_diag(prob, 2).items() if k != "q")
    assert np.any(oracle_diag(prob, 3)["kener"] != 0)
    # the last step computes whatever the frequency
    prob = diag_problem(state_frequency=9999)
    assert np.all(oracle_diag(prob, 1, last_time_step=5)["kener"] == 0)
    assert np.any(oracle_diag(prob, 1, last_time_step=1)["kener"] != 0)
    # disabled: nothing is registered
    o = oracle.OracleSession(build_problem(ne=2, nlev=26, qsize=3, seed=4, state_frequency=1))
    assert o.diagnostics() is None
    o.close()


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("nsteps,cpstar", [(1, False), (3, False), (2, True)])
def test_diagnostics_match_oracle(nsteps, cpstar):
    from hommexx_amd.hip import HipSession
    kw = dict(use_cpstar=True, moisture=True) if cpstar else {}
    prob = build_problem(ne=4, nlev=26, qsize=4, seed=5, state_frequency=1, disable_diagnostics=False, **kw)
    print("case: ne4 L26 Q4, %d steps, use_cpstar %s" % (nsteps, cpstar), flush=True)
    ref = oracle_diag(prob, nsteps)
    with HipSession(prob) as h:
        h.run(nsteps)
        out = h.diagnostics()
    for k in ref:
        a, b = out[k], ref[k]
        scale = max(np.abs(b).max(), 1e-300)
        err = np.abs(a - b).max() / scale
        print("%-9s max rel diff %.2e" % (k, err))
        # the first step's values before the advance come from the shared
        # initial state and are summed in the same order: bit for bit
        if nsteps == 1 and k in SLOTS_SET:
            before = [s for s in SLOTS_SET[k] if s != 1]
            np.testing.assert_array_equal(a[:, before], b[:, before], err_msg=k)
        assert err <= 1e-12, (k, err)
        if k in SLOTS_SET:
            unset = [s for s in range(4) if s not in SLOTS_SET[k]]
            assert np.all(a[:, unset] == 0), k
