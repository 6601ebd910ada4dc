"""hommexx_hip_adjoint_step_noalloc (the adjoint step on the library's fixed
device scratch array, for callers that cannot allocate) against
hommexx_hip_adjoint_step through the C-ABI: bit for bit, on two calls in a
row on the same session and then on a smaller configuration, so the second
configuration runs on an array that still holds the first one's contents."""
import ctypes

import pytest
import torch

from hommexx_amd.adjoint import DiffSession, initial_state
from hommexx_amd.problem import build_problem

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p
_D = ctypes.c_double


def both(prob, seed):
    with DiffSession(prob) as s:
        L = s.L
        L.hommexx_hip_adjoint_step_noalloc.argtypes = [_D, _P, _P, _P, _P, _P, _P]
        st = initial_state(prob)
        gen = torch.Generator(device="cuda").manual_seed(seed)
        lam = [torch.randn(a.shape, generator=gen, dtype=torch.float64, device="cuda") for a in st]
        for call in range(2):
            a = [t.clone() for t in lam]
            b = [t.clone() for t in lam]
            # the no-allocation step first: on the first call of a session
            # nothing of the allocating step exists yet
            torch.cuda.synchronize()
            rc = L.hommexx_hip_adjoint_step_noalloc(prob.dt, *[_P(t.data_ptr()) for t in st],
                                                    *[_P(t.data_ptr()) for t in b])
            if rc:
                s.h._check("hommexx_hip_adjoint_step_noalloc")
            assert rc == 0
            s.adjoint_step(st, a)
            for x, y in zip(a, b):
                assert torch.equal(x, y), (prob.nelem, prob.nlev, call)
            lam = a  # the next call starts from another adjoint


@pytest.mark.parametrize("ne,nlev,rsplit", [(4, 72, 1), (3, 26, 0), (2, 26, 3)])
def test_noalloc_adjoint_matches_allocating_step(ne, nlev, rsplit):
    """ne4 L72 fills most of the 128 MiB array; the next cases are smaller"""
    print("case: ne%d L%d rsplit %d" % (ne, nlev, rsplit), flush=True)
    prob = build_problem(ne=ne, nlev=nlev, qsize=2, seed=50 + ne, rsplit=rsplit, t_noise=0.1)
    both(prob, ne)


def test_noalloc_rejects_large_configurations():
    prob = build_problem(ne=8, nlev=72, qsize=0, seed=55)
    with DiffSession(prob) as s:
        L = s.L
        L.hommexx_hip_adjoint_step_noalloc.argtypes = [_D, _P, _P, _P, _P, _P, _P]
        st = initial_state(prob)
        lam = [torch.zeros_like(t) for t in st]
        rc = L.hommexx_hip_adjoint_step_noalloc(prob.dt, *[_P(t.data_ptr()) for t in st],
                                                *[_P(t.data_ptr()) for t in lam])
        assert rc == 3
        assert L.hommexx_hip_last_error() == 71
        L.hommexx_hip_clear_error()
