"""The pure-PyTorch FP64 step (the adjoint's reference gradient) against the
CPU oracle: the same dry remap interval (rsplit 0; rsplit 1 and 3 with the
PPM remap of the states, algs 1 and 2; qsplit 2) on the same seeded inputs."""
import pytest
import torch

from hommexx_amd.problem import build_problem
from oracle import OracleSession
from parity import rel_l2
from torch_ref import TorchStep, initial_state


@pytest.mark.parametrize("rsplit,alg,nu_top,nsteps,qsplit", [(0, 1, 0.0, 1, 1), (0, 1, 2.5e5, 3, 1),
                                                            (1, 1, 2.5e5, 3, 1), (1, 2, 2.5e5, 2, 1),
                                                            (3, 1, 2.5e5, 2, 1), (1, 1, 2.5e5, 2, 2),
                                                            (0, 1, 2.5e5, 2, 2)])
def test_torch_step_matches_oracle(rsplit, alg, nu_top, nsteps, qsplit):
    prob = build_problem(ne=4, nlev=26, qsize=0, seed=3, rsplit=rsplit, remap_alg=alg, nu_top=nu_top, t_noise=0.1,
                         qsplit=qsplit)
    o = OracleSession(prob)
    _, _, n0, _ = o.run(nsteps)
    ref = o.state()
    o.close()
    with torch.no_grad():
        v, T, ps = TorchStep(prob).run(*initial_state(prob), nsteps)
    for name, a, b in (("v", v, ref["v"][:, n0]), ("T", T, ref["T"][:, n0]), ("ps_v", ps, ref["ps_v"][:, n0])):
        e = rel_l2(a.numpy(), b)
        assert e <= 1e-13, (name, e)


def test_torch_step_rejects_other_configurations():
    with pytest.raises(ValueError):
        TorchStep(build_problem(ne=2, nlev=26, qsize=0, hypervis_scaling=3.2))
    with pytest.raises(ValueError):
        TorchStep(build_problem(ne=2, nlev=26, qsize=1, moisture=True))
