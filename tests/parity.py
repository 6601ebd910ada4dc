"""Shared helpers for comparing the HIP step with the CPU oracle."""
import numpy as np

FIELDS = ("v", "T", "dp3d", "Qdp", "Q", "ps_v", "omega_p")


def rel_l2(a, b):
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return d / nb if nb > 0 else d


def compare(gpu, cpu, fields=FIELDS):
    """relative L2 error and bit-exact fraction per field"""
    out = {}
    for k in fields:
        a, b = gpu[k], cpu[k]
        out[k] = (rel_l2(a, b), float(np.mean(a == b)))
    return out


def fmt(rep):
    return ", ".join("%s: %.2e (%.0f%% bitwise)" % (k, e, 100 * f) for k, (e, f) in rep.items())
