"""Differentiable form: the HIP adjoint (``hommexx_hip_adjoint_step`` behind
``torch.autograd.Function`` with checkpointing) against

* a pure-PyTorch FP64 reference gradient (``tests/torch_ref.py`` through
  ``torch.autograd``, on the CPU) at ne4 L26, and
* central finite differences of the HIP forward step, with a relative step
  of 1e-6, along 64 seeded directions at ne16 L72 Q4 over 48 steps with a
  checkpoint every 8 steps.

Dry configurations: rsplit = 0 (Eulerian vertical coordinate), rsplit = 1
(states remapped every step: the backward pass runs the linearised PPM remap
with the forward pass's branches frozen), rsplit = 3 (three dynamics steps
per remap) and qsplit = 2.

The gradient is taken of a regional kinetic-energy-weighted mismatch against
a seeded synthetic target, with respect to the initial T and ps (and v).
"""
import numpy as np
import pytest
import torch

from hommexx_amd.adjoint import (DiffSession, energy_mismatch, initial_state, region_weight, synthetic_target,
                                 trajectory)
from hommexx_amd.problem import build_problem
from parity import rel_l2
from torch_ref import TorchStep

pytestmark = pytest.mark.gpu


def ne16_problem(rsplit):
    # ne16 hyperviscosity as the reference's ne16 setups use it (nu = 7e15)
    nu = 7e15
    return build_problem(ne=16, nlev=72, qsize=4, seed=32, rsplit=rsplit, nu=nu, nu_p=nu, nu_q=nu, nu_s=nu,
                         nu_div=nu)


@pytest.mark.parametrize("rsplit,alg,qsplit", [(0, 1, 1), (1, 1, 1), (1, 2, 1), (3, 1, 1), (1, 1, 2), (0, 1, 2)])
def test_adjoint_matches_torch_reference_ne4(rsplit, alg, qsplit):
    prob = build_problem(ne=4, nlev=26, qsize=2, seed=31, rsplit=rsplit, remap_alg=alg, t_noise=0.1, qsplit=qsplit)
    nsteps, every = 6, 4
    print("case: ne4 L26 rsplit %d alg %d qsplit %d, %d steps, checkpoint every %d" % (rsplit, alg, qsplit, nsteps,
                                                                                       every), flush=True)
    # pure-PyTorch FP64 reference on the CPU
    x0 = [a.clone().requires_grad_() for a in initial_state(prob, "cpu")]
    out = TorchStep(prob).run(*x0, nsteps)
    Jr = energy_mismatch(prob, *out, synthetic_target(prob, 5, "cpu"), region_weight(prob, device="cpu"))
    gref = torch.autograd.grad(Jr, x0)
    # HIP forward + adjoint
    with DiffSession(prob) as s:
        x = [a.clone().requires_grad_() for a in initial_state(prob)]
        J = energy_mismatch(prob, *trajectory(s, *x, nsteps, every), synthetic_target(prob, 5), region_weight(prob))
        g = torch.autograd.grad(J, x)
    print("J: HIP %.17e torch %.17e" % (J.item(), Jr.item()))
    assert abs(J.item() - Jr.item()) <= 1e-11 * abs(Jr.item())
    for a, b, n in zip(g, gref, ("v", "T", "ps")):
        e = rel_l2(a.cpu().numpy(), b.numpy())
        print("grad %-2s rel L2 %.2e" % (n, e))
        assert e <= 1e-9, (n, e)


@pytest.mark.parametrize("rsplit", [0, 1, 3])
def test_dynamics_step_is_the_forward_step(rsplit):
    """hommexx_hip_dynamics_step (the backward pass's replay, no tracer step)
    gives the (v, T, ps) of prim_run_subcycle_c bit for bit in a dry run with
    tracers, and leaves the session's state alone."""
    prob = build_problem(ne=4, nlev=26, qsize=3, seed=34, rsplit=rsplit, t_noise=0.1)
    with DiffSession(prob) as s:
        for n in range(3):
            x = s.get_state()
            y = tuple(t.clone() for t in x)
            s.dynamics_step(y)
            assert all(torch.equal(a, b) for a, b in zip(s.get_state(), x))
            s.step()
            z = s.get_state()
            for a, b in zip(y, z):
                assert torch.equal(a, b), (rsplit, n)


@pytest.mark.parametrize("rsplit,qsplit", [(0, 1), (1, 1), (3, 1), (1, 2)])
def test_adjoint_segment_is_the_chained_step(rsplit, qsplit):
    """hommexx_hip_adjoint_segment over 4 steps (stage states kept) gives bit
    for bit what hommexx_hip_adjoint_step gives step by step on the states
    that hommexx_hip_dynamics_step replays."""
    prob = build_problem(ne=4, nlev=26, qsize=2, seed=35, rsplit=rsplit, qsplit=qsplit, t_noise=0.1)
    with DiffSession(prob) as s:
        x0 = initial_state(prob)
        states = [x0]
        for _ in range(3):
            y = tuple(t.clone() for t in states[-1])
            s.dynamics_step(y)
            states.append(y)
        gen = torch.Generator(device="cuda").manual_seed(7)
        lam = [torch.randn(a.shape, generator=gen, dtype=torch.float64, device="cuda") for a in x0]
        a = [t.clone() for t in lam]
        for st in reversed(states):
            s.adjoint_step(st, a)
        b = [t.clone() for t in lam]
        s.adjoint_segment(x0, 4, b)
    for p, q in zip(a, b):
        assert torch.equal(p, q), (rsplit, qsplit)


@pytest.mark.parametrize("rsplit", [0, 1, 3])
def test_adjoint_checkpoint_interval_independent(rsplit):
    """The gradient does not depend on the checkpoint interval (every = 1
    takes each step's input from the forward pass's checkpoints; 3 and 5
    replay the others with the dynamics-only step, tracers present), and the
    backward pass leaves the session at the forward pass's final state."""
    prob = build_problem(ne=4, nlev=26, qsize=3, seed=33, rsplit=rsplit)
    tgt, w = synthetic_target(prob, 2), region_weight(prob)
    grads = []
    with DiffSession(prob) as s:
        for every in (1, 3, 5):
            x = [a.clone().requires_grad_() for a in initial_state(prob)]
            out = trajectory(s, *x, 5, every)
            J = energy_mismatch(prob, *out, tgt, w)
            grads.append(torch.autograd.grad(J, x))
            for a, b in zip(s.get_state(), out):
                assert torch.equal(a, b)
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            assert torch.equal(a, b)


@pytest.mark.parametrize("rsplit", [0, 1, 3])
def test_adjoint_central_differences_ne16(rsplit):
    prob = ne16_problem(rsplit)
    # 48 dynamics steps: with rsplit 3 that is 16 remap intervals, checkpointed every 4
    nsteps, every, ndir, eps = (16, 4, 64, 1e-6) if rsplit == 3 else (48, 8, 64, 1e-6)
    print("case: ne16 L72 Q4 rsplit %d, %d steps, checkpoint every %d, %d directions" % (rsplit, nsteps, every, ndir),
          flush=True)
    tgt, w = synthetic_target(prob, 6), region_weight(prob)
    with DiffSession(prob) as s:
        v0, T0, ps0 = initial_state(prob)
        x = [v0.clone(), T0.clone().requires_grad_(), ps0.clone().requires_grad_()]
        J = energy_mismatch(prob, *trajectory(s, *x, nsteps, every), tgt, w)
        gT, gps = torch.autograd.grad(J, (x[1], x[2]))

        def Jof(T, ps):
            with torch.no_grad():
                return energy_mismatch(prob, *trajectory(s, v0, T, ps, nsteps, every), tgt, w).item()

        rng = np.random.default_rng(7)
        m = prob.mesh
        basis = np.stack([np.ones_like(m.lat), m.pos[..., 0], m.pos[..., 1], m.pos[..., 2],
                          m.pos[..., 0] * m.pos[..., 1], np.sin(3 * m.lon) * np.cos(m.lat)])
        to = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")
        worst = 0.0
        for n in range(ndir):
            # smooth random pattern plus white noise, relative to the state
            pat = np.tensordot(rng.standard_normal(len(basis)), basis, 1)
            prof = rng.standard_normal(prob.nlev)[None, :, None, None]
            dT = pat[:, None] * prof + 0.3 * rng.standard_normal(T0.shape)
            dps = np.tensordot(rng.standard_normal(len(basis)), basis, 1) + 0.3 * rng.standard_normal(ps0.shape)
            dT = to(dT) * T0 * eps
            dps = to(dps) * ps0 * eps
            fd = (Jof(T0 + dT, ps0 + dps) - Jof(T0 - dT, ps0 - dps)) / 2.0
            ad = (gT * dT).sum().item() + (gps * dps).sum().item()
            rel = abs(fd - ad) / abs(ad)
            worst = max(worst, rel)
            if n < 4 or rel > 1e-6:
                print("dir %2d  adjoint %.10e  FD %.10e  rel %.2e" % (n, ad, fd, rel), flush=True)
        print("worst relative difference over %d directions: %.2e" % (ndir, worst))
        assert worst <= 1e-5
