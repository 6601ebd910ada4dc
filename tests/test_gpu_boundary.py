"""The drop-in boundary on the GPU: the forcing pushes of the CAM path
(cxx_f90_interface.cpp:157-205) and a driver that keeps the library's build
dimensions (PREQX_PLEV, QSIZE_D) instead of calling hommexx_hip_configure."""
import ctypes

import numpy as np
import pytest

import oracle
from hommexx_amd.hip import HipSession
from hommexx_amd.problem import build_problem, make_forcing
from parity import compare, fmt

pytestmark = pytest.mark.gpu

_P = ctypes.POINTER(ctypes.c_double)


def ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def active(st, qsize):
    """the state without the tracer slots beyond qsize (the oracle keeps
    only the advanced tracers; the carried slots are checked separately)"""
    out = dict(st)
    out["Qdp"] = np.ascontiguousarray(st["Qdp"][:, :, :qsize])
    out["Q"] = np.ascontiguousarray(st["Q"][:, :qsize])
    return out


def seeded_tail(prob, seed):
    """non-zero data in the Qdp slots beyond qsize (never advanced)"""
    rng = np.random.default_rng(seed)
    q = prob.qsize
    prob.state["Qdp"][:, :, q:] = rng.uniform(1.0, 2.0, prob.state["Qdp"][:, :, q:].shape)


@pytest.mark.parametrize("ftype", [0, 2])
def test_forcing_pushes(ftype):
    """cxx_push_forcing_to_f90 returns what f90_push_forcing_to_cxx sent (FQ
    only with ftype 0, as in the reference), and f90_push_forcing_to_cxx
    copies the device Qdp, tail slots included, into the Fortran Qdp array
    (tracers.push_qdp, cxx_f90_interface.cpp:201) without changing the
    device state."""
    prob = build_problem(ne=4, nlev=26, qsize=3, qsize_d=5, seed=40 + ftype, ftype=ftype)
    seeded_tail(prob, 40)
    prob.forcing = make_forcing(prob, seed=41, fq_scale=3e-5)
    f = prob.forcing
    print("case: ne4 L26 Q3 qsize_d 5, ftype %d" % ftype, flush=True)
    o = oracle.OracleSession(prob)
    with HipSession(prob) as h:
        L = h.L
        FM = np.zeros_like(f["FM"])
        FT = np.zeros_like(f["FT"])
        FQ = np.full_like(f["FQ"], -7.0)
        L.cxx_push_forcing_to_f90(ptr(FM), ptr(FT), ptr(FQ))
        h._check("cxx_push_forcing_to_f90")
        assert np.array_equal(FM, f["FM"])
        assert np.array_equal(FT, f["FT"])
        if ftype == 0:
            assert np.array_equal(FQ, f["FQ"])
        else:
            assert np.all(FQ == -7.0)  # FQ is only exchanged with ftype 0
        # a Fortran Qdp that differs from the device state
        qf = np.random.default_rng(42).uniform(size=prob.state["Qdp"].shape)
        fa = [np.ascontiguousarray(f["FM"]), np.ascontiguousarray(f["FT"]), np.ascontiguousarray(f["FQ"])]
        L.f90_push_forcing_to_cxx(*[ptr(a) for a in fa], ptr(qf))
        h._check("f90_push_forcing_to_cxx")
        assert np.array_equal(qf, prob.state["Qdp"])
        h.run(2)
        out = h.state()
        L.f90_push_forcing_to_cxx(*[ptr(a) for a in fa], ptr(qf))
        assert np.array_equal(qf, out["Qdp"])
    o.run(2)
    ref = o.state()
    o.close()
    rep = compare(active(out, 3), active(ref, 3))
    print(fmt(rep))
    for k, (e, _) in rep.items():
        assert e <= 1e-12, (k, e)
    assert np.array_equal(out["Qdp"][:, :, 3:], prob.state["Qdp"][:, :, 3:])
    assert not np.any(out["Q"][:, 3:])


def test_build_dimensions_without_configure():
    """An unchanged Fortran driver: no hommexx_hip_configure call, so the
    library uses its build's PREQX_PLEV = 72 and QSIZE_D = 4 (the reference's
    defaults); 3 tracers are advanced and the fourth slot is carried."""
    prob = build_problem(ne=4, nlev=72, qsize=3, qsize_d=4, seed=44)
    seeded_tail(prob, 44)
    o = oracle.OracleSession(prob)
    o.run(2)
    ref = o.state()
    o.close()
    with HipSession(prob, configure=False) as h:
        h.run(2)
        out = h.state()
    rep = compare(active(out, 3), active(ref, 3))
    print(fmt(rep))
    for k, (e, _) in rep.items():
        assert e <= 1e-12, (k, e)
    assert np.array_equal(out["Qdp"][:, :, 3:], prob.state["Qdp"][:, :, 3:])
    assert not np.any(out["Q"][:, 3:])


def test_phase_timers():
    """hommexx_hip_profile_only restricts the HIP-event timers to one phase
    (the timed steps of bench.py carry only those events); without it every
    phase of the step is timed.  The timers do not change the results."""
    prob = build_problem(ne=2, nlev=26, qsize=4, seed=45)
    with HipSession(prob) as h:
        h.profile(True, only="euler_advect")
        h.run(2)
        h.profile(False)
        one = h.profile_report()
        h.profile_reset()
        h.profile(True)
        h.run(1)
        h.profile(False)
        every = h.profile_report()
        timed = h.state()
    assert set(one) == {"euler_advect"}
    assert one["euler_advect"][0] == 2 * 3 and one["euler_advect"][1] > 0  # three stages per step
    assert {"caar", "caar_dss", "hv_cycle", "euler_advect", "euler_dss", "remap"} <= set(every)
    with HipSession(prob) as h:
        h.run(3)
        plain = h.state()
    for k in ("v", "T", "dp3d", "Qdp"):
        assert np.array_equal(timed[k], plain[k]), k


def test_run_steps_driver(capsys):
    """python -m hommexx_amd.run_steps, at a small size, in both forms"""
    import sys
    from hommexx_amd import run_steps
    for form in ("fp64", "mixed"):
        argv = sys.argv
        sys.argv = ["run_steps", "--form", form, "--ne", "2", "--nlev", "26", "--qsize", "4", "--steps", "2",
                    "--warmup", "1"]
        try:
            run_steps.main()
        finally:
            sys.argv = argv
        assert ("%s ne2 L26 Q4: 2 steps" % form) in capsys.readouterr().out
