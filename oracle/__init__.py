"""ctypes binding of the CPU oracle (``oracle/preqx_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
(as the checker) and by ``bench.py``'s CPU-baseline leg.  The product
library never links, loads or calls anything here.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


def build(force=False):
    """Compile the oracle with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def _bind(path):
    """Load one build of the oracle and declare its entry points."""
    L = ctypes.CDLL(path)
    L.oracle_create.restype = _vp
    L.oracle_create.argtypes = [ctypes.c_int] * 4
    L.oracle_destroy.argtypes = [_vp]
    L.oracle_set_params.restype = ctypes.c_int
    L.oracle_set_params.argtypes = [_vp] + [ctypes.c_int] * 7 + [ctypes.c_double] * 6 + \
        [ctypes.c_int] * 2 + [ctypes.c_double] + [ctypes.c_int] * 2
    L.oracle_set_dvv.argtypes = [_vp, _dp]
    L.oracle_set_2d.argtypes = [_vp] + [_dp] * 11
    L.oracle_set_hvcoord.argtypes = [_vp, ctypes.c_double, _dp, _dp]
    L.oracle_add_connection.restype = ctypes.c_int
    L.oracle_add_connection.argtypes = [_vp] + [ctypes.c_int] * 8
    L.oracle_set_states.argtypes = [_vp] + [_dp] * 5
    L.oracle_get_states.argtypes = [_vp] + [_dp] * 7
    L.oracle_set_time_level.argtypes = [_vp] + [ctypes.c_int] * 5
    L.oracle_run_subcycle.restype = ctypes.c_int
    L.oracle_run_subcycle.argtypes = [_vp, ctypes.c_double, _ip, _ip, _ip, _ip]
    L.oracle_unit_gradient.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    L.oracle_unit_divergence.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.oracle_unit_vorticity.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.oracle_unit_laplace.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.oracle_unit_vlaplace.argtypes = [ctypes.c_int, ctypes.c_double] + [_dp] * 9
    L.oracle_unit_laplace_tensor.argtypes = [ctypes.c_int] + [_dp] * 6
    L.oracle_unit_vlaplace_cartesian.argtypes = [ctypes.c_int] + [_dp] * 7
    L.oracle_unit_limiter.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.oracle_unit_ppm.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
    L.oracle_unit_dss.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
    L.oracle_num_threads.restype = ctypes.c_int
    L.oracle_debug_stop.argtypes = [_vp, ctypes.c_int]
    L.oracle_set_forcing.argtypes = [_vp, _dp, _dp, _dp]
    L.oracle_set_diagnostics.argtypes = [_vp, ctypes.c_int, ctypes.c_int] + [_dp] * 8
    L.oracle_set_last_time_step.argtypes = [_vp, ctypes.c_int]
    return L


def lib():
    """The oracle (default build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        _lib = _bind(LIB_PATH)
    return _lib


_plants = {}


def plant(n):
    """The oracle built with planted mistake ``n`` (see ORACLE_PLANT in
    preqx_oracle.c); used only to show that the pinning tests catch it."""
    if n not in _plants:
        path = os.path.join(_HERE, "plants", "liboracle_plant%d.so" % n)
        subprocess.check_call(["make", "-s", "-C", _HERE, os.path.join("plants", os.path.basename(path))])
        _plants[n] = _bind(path)
    return _plants[n]


def ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleSession:
    """Runs the reference step on the CPU for a :class:`hommexx_amd.problem.Problem`."""

    def __init__(self, prob, L=None):
        L = lib() if L is None else L
        self.L = L
        self.prob = prob
        p = prob.params
        self.h = L.oracle_create(prob.nelem, prob.nlev, prob.qsize, prob.qsize_d)
        rc = L.oracle_set_params(self.h, p.remap_alg, p.limiter_option, p.rsplit, p.qsplit, p.time_step_type,
                                 p.qsize, p.state_frequency, p.nu, p.nu_p, p.nu_q, p.nu_s, p.nu_div, p.nu_top,
                                 p.hypervis_order, p.hypervis_subcycle, p.hypervis_scaling, p.ftype,
                                 int(p.moisture))
        if rc:
            raise ValueError("oracle rejected parameters (code %d)" % rc)
        self._keep = {k: f64(v) for k, v in prob.metric.items()}
        m = self._keep
        L.oracle_set_dvv(self.h, ptr(f64(prob.dvv)))
        L.oracle_set_2d(self.h, ptr(m["D"]), ptr(m["Dinv"]), ptr(m["fcor"]), ptr(m["mp"]), ptr(m["spheremp"]),
                        ptr(m["rspheremp"]), ptr(m["metdet"]), ptr(m["metinv"]), ptr(f64(prob.state["phis"])),
                        ptr(m["tensorvisc"]), ptr(m["vec_sph2cart"]))
        L.oracle_set_hvcoord(self.h, prob.vcoord.ps0, ptr(f64(prob.vcoord.hyai)), ptr(f64(prob.vcoord.hybi)))
        for c in prob.connections:
            rc = L.oracle_add_connection(self.h, *c)
            if rc:
                raise ValueError("bad connection %r" % (c,))
        s = prob.state
        L.oracle_set_states(self.h, ptr(f64(s["v"])), ptr(f64(s["T"])), ptr(f64(s["dp3d"])), ptr(f64(s["Qdp"])),
                            ptr(f64(s["ps_v"])))
        L.oracle_set_time_level(self.h, *prob.time_level)
        if getattr(prob, "forcing", None) is not None:
            f = prob.forcing
            self._keep["forcing"] = [f64(f["FM"]), f64(f["FT"]), f64(f["FQ"])]
            L.oracle_set_forcing(self.h, *[ptr(a) for a in self._keep["forcing"]])
        self.diag = None
        if not p.disable_diagnostics:
            from hommexx_amd.problem import DIAGNOSTIC_ARRAYS, diagnostic_arrays
            self.diag = diagnostic_arrays(prob)
            L.oracle_set_diagnostics(self.h, 0, int(p.use_cpstar), *[ptr(self.diag[k]) for k in DIAGNOSTIC_ARRAYS])
        self.tl = None

    def run(self, nsteps=1, dt=None, last_time_step=1 << 30):
        L = self.L
        dt = self.prob.dt if dt is None else dt
        L.oracle_set_last_time_step(self.h, last_time_step)
        a = [ctypes.c_int() for _ in range(4)]
        for _ in range(nsteps):
            rc = L.oracle_run_subcycle(self.h, dt, *[ctypes.byref(x) for x in a])
            if rc:
                raise RuntimeError("oracle step failed (code %d)" % rc)
        self.tl = tuple(x.value for x in a)  # nstep, nm1, n0, np1 (0-based)
        return self.tl

    def state(self):
        pr = self.prob
        s = pr.state
        out = {k: np.zeros_like(s[k]) for k in ("v", "T", "dp3d", "Qdp", "Q", "ps_v")}
        out["omega_p"] = np.zeros((pr.nelem, pr.nlev, 4, 4))
        self.L.oracle_get_states(self.h, ptr(out["v"]), ptr(out["T"]), ptr(out["dp3d"]), ptr(out["Qdp"]),
                                ptr(out["Q"]), ptr(out["ps_v"]), ptr(out["omega_p"]))
        return out

    def diagnostics(self):
        """Copies of the diagnostic arrays (None with disable_diagnostics)."""
        return None if self.diag is None else {k: v.copy() for k, v in self.diag.items()}

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
