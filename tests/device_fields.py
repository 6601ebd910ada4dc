"""Comparisons of two runs on the GPU without host copies of the state:
the fields of a run are taken with hommexx_hip_get_field into torch tensors
on the device (one 3-D field at a time where it matters), and the relative L2
differences and mass integrals are reduced there."""
import ctypes

import torch

F64 = torch.float64
# (name, which, components per point and level)
DYN = (("v", 0, 2), ("T", 1, 1), ("dp3d", 2, 1))


def _get(h, which, q, t):
    rc = h.L.hommexx_hip_get_field(ctypes.c_int(which), ctypes.c_int(q), ctypes.c_void_p(t.data_ptr()))
    if rc:
        h._check("hommexx_hip_get_field")
        raise RuntimeError("hommexx_hip_get_field(%d, %d) failed (%d)" % (which, q, rc))
    return t


def field(h, name, q=0):
    p = h.prob
    E, L = p.nelem, p.nlev
    shape = {"v": (E, L, 2, 16), "T": (E, L, 16), "dp3d": (E, L, 16), "ps_v": (E, 16), "Qdp": (E, L, 16),
             "Q": (E, L, 16)}[name]
    which = {"v": 0, "T": 1, "dp3d": 2, "ps_v": 3, "Qdp": 4, "Q": 5}[name]
    torch.cuda.synchronize()
    return _get(h, which, q, torch.empty(shape, dtype=F64, device="cuda"))


def snapshot(h, tracers=("Qdp",)):
    """v, T, dp3d, ps_v and the given per-tracer fields of the current state"""
    out = {n: field(h, n) for n in ("v", "T", "dp3d", "ps_v")}
    for n in tracers:
        out[n] = torch.stack([field(h, n, q) for q in range(h.prob.qsize)])
    return out


def rel_errors(h, ref):
    """relative L2 difference of every field of ``ref`` (a snapshot of
    another run) from the current state of session ``h``"""
    err = {}
    for n, r in ref.items():
        if n in ("Qdp", "Q"):
            num = den = 0.0
            for q in range(r.shape[0]):
                a = field(h, n, q)
                num += float(((a - r[q]) ** 2).sum())
                den += float((r[q] ** 2).sum())
                del a
            err[n] = (num / den) ** 0.5 if den > 0 else num ** 0.5
        else:
            a = field(h, n)
            err[n] = float(torch.linalg.vector_norm(a - r) / torch.linalg.vector_norm(r))
    return err


def masses(h):
    """dry mass sum(spheremp * ps) and per-tracer mass sum(spheremp * Qdp)"""
    p = h.prob
    sph = torch.as_tensor(p.metric["spheremp"].reshape(p.nelem, 16), dtype=F64, device="cuda")
    dry = float((sph * field(h, "ps_v")).sum())
    q = [float((sph[:, None, :] * field(h, "Qdp", i)).sum()) for i in range(p.qsize)]
    return dry, torch.tensor(q, dtype=F64)
