"""Host-side checks of the differentiable form that need no GPU: the
configurations the adjoint does not cover are rejected before the library
is touched, and the loss helpers behave as documented."""
import numpy as np
import pytest
import torch

from hommexx_amd.adjoint import DiffSession, energy_mismatch, initial_state, region_weight, synthetic_target
from hommexx_amd.hip import HommexxError
from hommexx_amd.problem import build_problem


@pytest.mark.parametrize("over", [dict(moisture=True), dict(ftype=0), dict(hypervis_scaling=3.2)])
def test_unsupported_configurations_are_rejected(over):
    prob = build_problem(ne=2, nlev=26, qsize=1, **over)
    with pytest.raises(HommexxError):
        DiffSession(prob)


def test_energy_mismatch_is_zero_at_the_target_and_quadratic():
    prob = build_problem(ne=2, nlev=26, qsize=0)
    tgt = synthetic_target(prob, 3, device="cpu")
    w = region_weight(prob, device="cpu")
    assert energy_mismatch(prob, *tgt, tgt, w).item() == 0.0
    v, T, ps = initial_state(prob, "cpu")
    j1 = energy_mismatch(prob, v, T, ps, tgt, w).item()
    # scaling the mismatch by 2 scales the loss by 4
    d = [2.0 * a - b for a, b in zip((v, T, ps), tgt)]  # target + 2 * (x - target)
    j2 = energy_mismatch(prob, *d, tgt, w).item()
    assert j1 > 0 and abs(j2 / j1 - 4.0) < 1e-12


def test_region_weight_peaks_at_the_centre():
    prob = build_problem(ne=4, nlev=26, qsize=0)
    w = region_weight(prob, lat0=40.0, lon0=60.0, device="cpu").numpy() / prob.metric["spheremp"]
    m = prob.mesh
    i = np.unravel_index(np.argmax(w), w.shape)
    assert abs(np.degrees(m.lat[i]) - 40.0) < 10.0 and abs(np.degrees(m.lon[i]) - 60.0) < 15.0
    assert w.max() <= 1.0 and w.min() >= 0.0
