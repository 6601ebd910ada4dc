"""Differentiable form of the preqx step: a ``torch.autograd.Function`` over
the C-ABI.

The forward pass runs ``prim_run_subcycle_c`` (the HIP step) and keeps the
prognostic state (v, T, ps) every ``every`` steps as device checkpoints.  The
backward pass hands each segment, last first, to
``hommexx_hip_adjoint_segment``: from the segment's checkpoint it recomputes
the dynamics of the segment's steps once (the same map without the tracer
step, which a dry run's dynamics do not see), keeping every stage state, and
runs the hand-written HIP adjoint kernels of CAAR, hyperviscosity, DSS and
the PPM remap over the steps in reverse.  The session's model state is not
touched and stays at the forward pass's end.  Device memory:
``ceil(nsteps/every)`` checkpoints (v, T, ps) plus the stage states of one
segment inside the library.

The adjoint covers dry configurations with any ``rsplit`` and ``qsplit``: one
``prim_run_subcycle_c`` call is qsplit * max(rsplit, 1) dynamics steps (RK5
stages and hyperviscosity) followed by the remap.  With ``rsplit > 0`` the
states are remapped and the backward pass runs the linearised PPM remap
(branches frozen from the forward pass); with ``rsplit = 0`` the remap moves
only tracers.  Tracers do not act on the dynamics of a dry run, so they have
no adjoint.

State tensors use the library's device layouts: v ``(nelem, nlev, 2, 4, 4)``,
T ``(nelem, nlev, 4, 4)``, ps ``(nelem, 4, 4)``, float64 on the GPU.
"""
import ctypes
import math

import numpy as np
import torch

from .hip import HipSession, HommexxError

_P = ctypes.c_void_p
_D = ctypes.c_double


def _ptr(t):
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return _P(t.data_ptr())


class DiffSession:
    """A :class:`HipSession` with device-side state access and the adjoint
    step.  ``prob.params`` must describe a dry run with constant-coefficient
    hyperviscosity and no forcing."""

    def __init__(self, prob, device="cuda"):
        p = prob.params
        if p.moisture or p.hypervis_scaling != 0.0 or p.ftype != -1:
            raise HommexxError("the adjoint needs no moisture, constant HV and no forcing")
        self.prob = prob
        self.device = torch.device(device)
        self.h = HipSession(prob)
        L = self.h.L
        L.hommexx_hip_get_state.argtypes = [_P, _P, _P]
        L.hommexx_hip_set_state.argtypes = [_P, _P, _P]
        L.hommexx_hip_adjoint_step.argtypes = [_D, _P, _P, _P, _P, _P, _P]
        L.hommexx_hip_dynamics_step.argtypes = [_D, _P, _P, _P]
        L.hommexx_hip_adjoint_segment.argtypes = [_D, ctypes.c_int, _P, _P, _P, _P, _P, _P]
        self.L = L
        E, nl = prob.nelem, prob.nlev
        self.shapes = ((E, nl, 2, 4, 4), (E, nl, 4, 4), (E, 4, 4))

    def empty(self):
        return tuple(torch.empty(s, dtype=torch.float64, device=self.device) for s in self.shapes)

    def get_state(self):
        out = self.empty()
        torch.cuda.synchronize(self.device)
        if self.L.hommexx_hip_get_state(*[_ptr(t) for t in out]):
            self.h._check("hommexx_hip_get_state")
        return out

    def set_state(self, v, T, ps):
        args = [x.contiguous() for x in (v, T, ps)]
        torch.cuda.synchronize(self.device)
        if self.L.hommexx_hip_set_state(*[_ptr(t) for t in args]):
            self.h._check("hommexx_hip_set_state")

    def step(self, nsteps=1):
        self.h.run(nsteps)

    def dynamics_step(self, state):
        """(v, T, ps) -> its image under one step's dynamics (and remap), in
        place; the session's model state is not changed"""
        torch.cuda.synchronize(self.device)
        rc = self.L.hommexx_hip_dynamics_step(self.prob.dt, *[_ptr(t) for t in state])
        if rc:
            self.h._check("hommexx_hip_dynamics_step")
            raise HommexxError("hommexx_hip_dynamics_step failed (%d)" % rc)

    def adjoint_step(self, state, lam):
        """lam (adjoint of the step's output) -> adjoint of its input, in place"""
        torch.cuda.synchronize(self.device)
        rc = self.L.hommexx_hip_adjoint_step(self.prob.dt, *[_ptr(t) for t in state], *[_ptr(t) for t in lam])
        if rc:
            self.h._check("hommexx_hip_adjoint_step")
            raise HommexxError("hommexx_hip_adjoint_step failed (%d)" % rc)

    def adjoint_segment(self, state, n, lam):
        """lam (adjoint of the state after n steps from ``state``) -> adjoint
        of ``state``, in place"""
        torch.cuda.synchronize(self.device)
        rc = self.L.hommexx_hip_adjoint_segment(self.prob.dt, n, *[_ptr(t) for t in state],
                                                *[_ptr(t) for t in lam])
        if rc:
            self.h._check("hommexx_hip_adjoint_segment")
            raise HommexxError("hommexx_hip_adjoint_segment failed (%d)" % rc)

    def close(self):
        self.h.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class PreqxTrajectory(torch.autograd.Function):
    """(v, T, ps) at step 0 -> (v, T, ps) after ``nsteps`` HIP steps."""

    @staticmethod
    def forward(ctx, v, T, ps, sess, nsteps, every):
        sess.set_state(v, T, ps)
        ckpts = []
        for n in range(nsteps):
            if n % every == 0:
                ckpts.append(sess.get_state())
            sess.step()
        out = sess.get_state()
        ctx.sess, ctx.nsteps, ctx.every, ctx.ckpts = sess, nsteps, every, ckpts
        return out

    @staticmethod
    def backward(ctx, gv, gT, gps):
        sess, nsteps, every = ctx.sess, ctx.nsteps, ctx.every
        lam = [torch.zeros(s, dtype=torch.float64, device=sess.device) if g is None else g.contiguous().clone()
               for g, s in zip((gv, gT, gps), sess.shapes)]
        for s in reversed(range(len(ctx.ckpts))):
            n0 = s * every
            n1 = min(n0 + every, nsteps)
            sess.adjoint_segment(ctx.ckpts[s], n1 - n0, lam)
        ctx.ckpts = None
        return lam[0], lam[1], lam[2], None, None, None


def trajectory(sess, v, T, ps, nsteps, every=8):
    return PreqxTrajectory.apply(v, T, ps, sess, nsteps, every)


# ------------------------------------------------------------------ loss
G = 9.80616
CP = 1005.0
RGAS = 287.04
T_REF = 300.0
P_REF = 1.0e5


def initial_state(prob, device="cuda"):
    """(v, T, ps) of time level n0 of ``prob``, device layout"""
    s = prob.state
    n0 = prob.time_level[1] - 1
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    return t(s["v"][:, n0]), t(s["T"][:, n0]), t(s["ps_v"][:, n0])


def synthetic_target(prob, seed=0, device="cuda"):
    """A seeded target state: the initial state plus smooth large-scale
    perturbations (a few m/s, about 1 K, a few hPa)."""
    rng = np.random.default_rng(seed)
    m = prob.mesh
    x, y, z = m.pos[..., 0], m.pos[..., 1], m.pos[..., 2]
    basis = np.stack([x, y, z, x * y, y * z, z * x, x * x - y * y])
    pat = lambda: np.tensordot(rng.standard_normal(len(basis)), basis, 1)
    prof = np.sin(np.linspace(0.2, 3.0, prob.nlev))[None, :, None, None]
    v, T, ps = (a.cpu().numpy() for a in initial_state(prob, "cpu"))
    v = v.copy()
    v[:, :, 0] += 3.0 * pat()[:, None] * prof
    v[:, :, 1] += 3.0 * pat()[:, None] * prof
    T = T + 1.0 * pat()[:, None] * prof
    ps = ps + 300.0 * pat()
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    return t(v), t(T), t(ps)


def region_weight(prob, lat0=40.0, lon0=60.0, radius=2.0e6, device="cuda"):
    """Gaussian regional mask around (lat0, lon0) times the GLL mass
    ``spheremp``, shape (nelem, 4, 4)."""
    m = prob.mesh
    la, lo = math.radians(lat0), math.radians(lon0)
    c = np.sin(la) * np.sin(m.lat) + np.cos(la) * np.cos(m.lat) * np.cos(m.lon - lo)
    d = 6.376e6 * np.arccos(np.clip(c, -1.0, 1.0))
    w = np.exp(-(d / radius) ** 2) * prob.metric["spheremp"]
    return torch.as_tensor(w, dtype=torch.float64, device=device)


def energy_mismatch(prob, v, T, ps, target, weight):
    """Regional kinetic-energy-weighted mismatch (total-energy norm):
    0.5 * sum w * dp_ref/g * (|v - v*|^2 + cp/T_ref (T - T*)^2)
    + 0.5 * sum w * R T_ref / p_ref^2 * (ps - ps*)^2 / g * p_ref"""
    vc = prob.vcoord
    dpref = torch.as_tensor((vc.hyai[1:] - vc.hyai[:-1]) * vc.ps0 + (vc.hybi[1:] - vc.hybi[:-1]) * vc.ps0,
                            dtype=torch.float64, device=v.device).view(1, -1, 1, 1)
    vt, Tt, pt = target
    wk = weight.unsqueeze(1) * dpref / G
    ke = ((v - vt) ** 2).sum(2)
    te = (CP / T_REF) * (T - Tt) ** 2
    pe = (RGAS * T_REF / P_REF ** 2) * (ps - pt) ** 2 * (P_REF / G)
    return 0.5 * (wk * (ke + te)).sum() + 0.5 * (weight * pe).sum()
