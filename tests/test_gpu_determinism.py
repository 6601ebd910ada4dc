"""Run-to-run reproducibility at the benchmark size (ne30, 72 levels, 40
tracers, dt 300 s, nu 1e15 subcycled 3 times, remap every step).

The step has no floating-point atomics and no reduction whose order depends
on scheduling: every sum runs in a fixed lane or loop order (the DSS in the
reference's accumulation order, the limiter sums serially over the 16
points, the column scans level by level).  Two runs from the same state must
therefore agree bit for bit, in both forms; a race between kernels or inside
one (a missing barrier, a buffer read while another kernel writes it) would
show here as a difference.
"""
import numpy as np
import pytest

from hommexx_amd.hip import HipSession
from hommexx_amd.problem import build_problem

pytestmark = pytest.mark.gpu

NU = 1e15


@pytest.fixture(scope="module")
def prob():
    return build_problem(ne=30, nlev=72, qsize=40, seed=1, dt=300.0, nu=NU, nu_p=NU, nu_q=NU, nu_s=NU, nu_div=NU,
                         nu_top=2.5e5, hypervis_subcycle=3, rsplit=1)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_two_runs_agree_bit_for_bit(prob, form):
    outs = []
    for _ in range(2):
        with HipSession(prob, abort_on_error=True, form=form) as h:
            h.run(3)
            outs.append(h.state())
    a, b = outs
    assert set(a) == set(b)
    for k in a:
        if isinstance(a[k], np.ndarray):
            same = np.array_equal(a[k], b[k], equal_nan=True)
            assert same, "%s: %d values differ between two %s runs" % (k, int((a[k] != b[k]).sum()), form)
    print("ne30 L72 Q40, %s form, 3 steps: two runs agree bit for bit (%s)" % (form, ", ".join(sorted(a))),
          flush=True)
