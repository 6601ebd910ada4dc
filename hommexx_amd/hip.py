"""ctypes harness over libhommexx_amd.so, driving the reference's C-ABI.

:class:`HipSession` replays the Fortran host's call sequence
(src/share/prim_driver_mod.F90:1042-1105 and prim_cxx_driver_mod.F90:131-139):
connectivity, derivative, simulation parameters, 2-D element data, states,
hybrid coordinates, boundary exchanges, time levels, then
``prim_run_subcycle_c`` per remap interval and ``cxx_push_results_to_f90``.
Arguments are passed the way Fortran passes them through ``bind(c)``:
scalars by reference, arrays as references to ``c_ptr`` values.

The HIP library is the only compute path: if it is not built this module
raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

from .problem import DIAGNOSTIC_ARRAYS, diagnostic_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhommexx_amd.so")

_I = ctypes.c_int
_D = ctypes.c_double
_B = ctypes.c_bool
_P = ctypes.c_void_p
_lib = None


class HommexxError(RuntimeError):
    pass


def load():
    """Load the shared library (built by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64.so.7: load it first, so the library
        # binds to the same HIP runtime (loaded after the library, torch would
        # start a second runtime that does not see the GPU)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise HommexxError("libhommexx_amd.so is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        L.hommexx_hip_error_message.restype = ctypes.c_char_p
        L.hommexx_hip_stream.restype = _P
        L.hommexx_hip_profile_get.argtypes = [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_D)]
        L.hommexx_hip_profile_names.argtypes = [ctypes.c_char_p, _I]
        L.hommexx_hip_configure.argtypes = [_I, _I]
        L.hommexx_hip_set_abort_on_error.argtypes = [_I]
        L.hommexx_hip_profile.argtypes = [_I]
        L.hommexx_hip_profile_only.argtypes = [ctypes.c_char_p]
        L.hommexx_hip_debug_stop.argtypes = [_I]
        L.hommexx_hip_set_form.argtypes = [_I]
        _lib = L
    return _lib


def _r(x, t):
    return ctypes.byref(t(x))


def _a(arr):
    """reference to a c_ptr holding the array address (Fortran type(c_ptr) by reference)"""
    return ctypes.byref(_P(arr.ctypes.data))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class HipSession:
    """One simulation on the GPU through the C-ABI.  The library keeps a
    single context per process (like the reference's Context singleton), so
    only one session may be alive at a time."""

    _alive = False

    FORMS = {"fp64": 0, "mixed": 1}

    def __init__(self, prob, abort_on_error=False, form="fp64", configure=True):
        """``configure=False`` leaves nlev and qsize_d at the library's build
        defaults (PREQX_PLEV, QSIZE_D), as an unchanged Fortran driver does;
        ``prob`` must then match them."""
        L = load()
        if HipSession._alive:
            raise HommexxError("a HipSession is already active; close() it first")
        self.prob = prob
        self.L = L
        L.hommexx_hip_set_abort_on_error(1 if abort_on_error else 0)
        if L.hommexx_hip_set_form(self.FORMS[form]):
            raise HommexxError("unknown form %r" % form)
        L.hommexx_hip_clear_error()
        L.initialize_hommexx_session()
        self._check("initialize_hommexx_session")
        HipSession._alive = True
        self._closed = False
        if configure:
            rc = L.hommexx_hip_configure(prob.nlev, prob.qsize_d)
            if rc:
                raise HommexxError("hommexx_hip_configure failed (%d)" % rc)
        p = prob.params
        # prim_init1: connectivity
        L.reset_cxx_comm(_r(0, _I))
        L.init_connectivity(_r(prob.nelem, _I))
        for c in prob.connections:
            L.add_connection(*[_r(x, _I) for x in c])
        L.finalize_connectivity()
        self._check("connectivity")
        # prim_init2
        self._keep = []
        dvv = f64(prob.dvv)
        self._keep.append(dvv)
        L.init_derivative_c(_a(dvv))
        L.init_simulation_params_c(
            _r(p.remap_alg, _I), _r(p.limiter_option, _I), _r(p.rsplit, _I), _r(p.qsplit, _I),
            _r(p.time_step_type, _I), _r(p.energy_fixer, _I), _r(p.qsize, _I), _r(p.state_frequency, _I),
            _r(p.nu, _D), _r(p.nu_p, _D), _r(p.nu_q, _D), _r(p.nu_s, _D), _r(p.nu_div, _D), _r(p.nu_top, _D),
            _r(p.hypervis_order, _I), _r(p.hypervis_subcycle, _I), _r(p.hypervis_scaling, _D), _r(p.ftype, _I),
            _r(p.prescribed_wind, _B), _r(p.moisture, _B), _r(p.disable_diagnostics, _B), _r(p.use_cpstar, _B),
            _r(p.use_semi_lagrangian_transport, _B))
        self._check("init_simulation_params_c")
        m = {k: f64(v) for k, v in prob.metric.items()}
        phis = f64(prob.state["phis"])
        self._keep += list(m.values()) + [phis]
        L.init_elements_2d_c(_r(prob.nelem, _I), _a(m["D"]), _a(m["Dinv"]), _a(m["fcor"]), _a(m["mp"]),
                             _a(m["spheremp"]), _a(m["rspheremp"]), _a(m["metdet"]), _a(m["metinv"]), _a(phis),
                             _a(m["tensorvisc"]), _a(m["vec_sph2cart"]), _r(p.hypervis_scaling == 0, _B))
        self._check("init_elements_2d_c")
        s = prob.state
        st = [f64(s[k]) for k in ("v", "T", "dp3d", "Qdp", "ps_v")]
        L.init_elements_states_c(*[_a(x) for x in st])
        self._check("init_elements_states_c")
        # init_diagnostics_c: the arrays prim_run_subcycle_c fills when
        # diagnostics are enabled (placeholders otherwise)
        self.diag = None
        if p.disable_diagnostics:
            diag = [np.zeros(1) for _ in range(8)]
        else:
            self.diag = diagnostic_arrays(prob)
            diag = [self.diag[k] for k in DIAGNOSTIC_ARRAYS]
        self._keep += diag
        L.init_diagnostics_c(*[_a(x) for x in diag])
        vc = prob.vcoord
        hy = [f64(vc.hyam), f64(vc.hyai), f64(vc.hybm), f64(vc.hybi)]
        self._keep += hy
        L.init_hvcoord_c(_r(vc.ps0, _D), *[_a(x) for x in hy])
        L.init_boundary_exchanges_c()
        L.init_time_level_c(*[_r(x, _I) for x in prob.time_level])
        self._check("init")
        if getattr(prob, "forcing", None) is not None:
            # prim_driver_mod.F90 pushes FM, FT, FQ (and Qdp) before the step
            f = prob.forcing
            fa = [f64(f["FM"]), f64(f["FT"]), f64(f["FQ"]), st[3]]
            self._keep += fa
            P = ctypes.POINTER(_D)
            L.f90_push_forcing_to_cxx(*[a.ctypes.data_as(P) for a in fa])
            self._check("f90_push_forcing_to_cxx")
        self.tl = None

    def _check(self, what):
        code = self.L.hommexx_hip_last_error()
        if code:
            msg = self.L.hommexx_hip_error_message().decode()
            self.L.hommexx_hip_clear_error()
            raise HommexxError("%s failed (%d): %s" % (what, code, msg))

    def run(self, nsteps=1, dt=None, last_time_step=1 << 30):
        dt = self.prob.dt if dt is None else dt
        nstep, nm1, n0, np1 = _I(), _I(), _I(), _I()
        last = _I(last_time_step)
        for _ in range(nsteps):
            self.L.prim_run_subcycle_c(ctypes.byref(_D(dt)), ctypes.byref(nstep), ctypes.byref(nm1),
                                       ctypes.byref(n0), ctypes.byref(np1), ctypes.byref(last))
            self._check("prim_run_subcycle_c")
        self.tl = (nstep.value, nm1.value, n0.value, np1.value)
        return self.tl

    def state(self):
        pr = self.prob
        s = pr.state
        out = {k: np.zeros_like(s[k]) for k in ("v", "T", "dp3d", "Qdp", "Q", "ps_v")}
        out["omega_p"] = np.zeros((pr.nelem, pr.nlev, 4, 4))
        self.L.cxx_push_results_to_f90(_a(out["v"]), _a(out["T"]), _a(out["dp3d"]), _a(out["Qdp"]),
                                       _a(out["Q"]), _a(out["ps_v"]), _a(out["omega_p"]))
        self._check("cxx_push_results_to_f90")
        return out

    def diagnostics(self):
        """Copies of the diagnostic arrays (None with disable_diagnostics)."""
        return None if self.diag is None else {k: v.copy() for k, v in self.diag.items()}

    def synchronize(self):
        self.L.hommexx_hip_synchronize()

    def profile(self, enable=True, only=None):
        """HIP-event timing of the step phases; only: time just that phase"""
        self.L.hommexx_hip_profile_only(only.encode() if only else None)
        self.L.hommexx_hip_profile(1 if enable else 0)

    def profile_reset(self):
        self.L.hommexx_hip_profile_reset()

    def profile_report(self):
        n = self.L.hommexx_hip_profile_names(None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        self.L.hommexx_hip_profile_names(buf, n + 1)
        out = {}
        for name in buf.value.decode().split("\n"):
            if not name:
                continue
            c, t = _I(), _D()
            self.L.hommexx_hip_profile_get(name.encode(), ctypes.byref(c), ctypes.byref(t))
            out[name] = (c.value, t.value)
        return out

    def close(self):
        if not getattr(self, "_closed", True):
            self.L.hommexx_hip_set_form(0)
            self.L.finalize_hommexx_session()
            self._closed = True
            HipSession._alive = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
