"""Host-side replay of the reference's ``prim_init2`` inputs.

A :class:`Problem` bundles everything the Fortran host hands to the C-ABI
before the first ``prim_run_subcycle_c`` (src/share/prim_driver_mod.F90:1042-1105):
simulation parameters, the derivative matrix, 2-D metric arrays, prognostic
states, hybrid coordinates, element connections and time levels.  All arrays
are in the flat Fortran layouts of the C-ABI.
"""
from dataclasses import dataclass, field

import numpy as np

from .ics import VCoord, jw_baroclinic
from .mesh import CubedSphere, derivative_matrix


@dataclass
class Params:
    """Arguments of ``init_simulation_params_c`` (cxx_f90_interface.cpp:33)."""
    remap_alg: int = 1
    limiter_option: int = 8
    rsplit: int = 1
    qsplit: int = 1
    time_step_type: int = 5
    energy_fixer: int = 0
    qsize: int = 4
    state_frequency: int = 9999
    nu: float = 1e16
    nu_p: float = 1e16
    nu_q: float = 1e16
    nu_s: float = 1e16
    nu_div: float = 1e16
    nu_top: float = 2.5e5
    hypervis_order: int = 2
    hypervis_subcycle: int = 3
    hypervis_scaling: float = 0.0
    ftype: int = -1
    prescribed_wind: bool = False
    moisture: bool = False
    disable_diagnostics: bool = True
    use_cpstar: bool = False
    use_semi_lagrangian_transport: bool = False


@dataclass
class Problem:
    ne: int
    nlev: int
    qsize: int
    dt: float
    params: Params
    mesh: CubedSphere = field(repr=False)
    vcoord: VCoord = field(repr=False)
    dvv: np.ndarray = field(repr=False)
    metric: dict = field(repr=False)
    state: dict = field(repr=False)
    connections: list = field(repr=False)
    time_level: tuple = (1, 2, 3, 0, 0)  # nm1, n0, np1 (1-based), nstep, nstep0
    # CAM forcing (ftype 0/2) as f90_push_forcing_to_cxx receives it:
    # {"FM": (nelem,nlev,2,4,4), "FT": (nelem,nlev,4,4), "FQ": (nelem,qsize_d,nlev,4,4)}
    forcing: dict = field(default=None, repr=False)

    @property
    def nelem(self):
        return self.mesh.nelem

    @property
    def qsize_d(self):
        return self.state["qsize_d"]


_MESH_CACHE = {}


def build_problem(ne=4, nlev=26, qsize=4, seed=0, dt=300.0, t_noise=0.0, u_perturb=1.0, qsize_d=None,
                  **param_overrides):
    """Synthetic JW-baroclinic problem on an ``ne`` cubed sphere.

    Defaults follow the small CPU configuration (ne=4, L26, Q4, dt=300 s,
    nu=1e16, hypervis subcycle 3, remap every step, limiter 8, remap alg 1,
    no forcing).  ``u_perturb`` scales the 1 m/s Gaussian jet perturbation
    (0 gives the steady balanced jet).  ``qsize_d`` is the tracer dimension
    of the Fortran arrays (QSIZE_D, default ``max(qsize, 1)``); slots beyond
    ``qsize`` hold seeded data that the step must carry unchanged.
    """
    if ne not in _MESH_CACHE:
        _MESH_CACHE[ne] = CubedSphere(ne)
    mesh = _MESH_CACHE[ne]
    vc = VCoord(nlev)
    p = Params(qsize=qsize, **param_overrides)
    state = jw_baroclinic(mesh, vc, qsize, seed=seed, qsize_d=qsize_d, t_noise=t_noise, u_perturb=u_perturb)
    return Problem(ne=ne, nlev=nlev, qsize=qsize, dt=dt, params=p, mesh=mesh, vcoord=vc,
                   dvv=np.ascontiguousarray(derivative_matrix()), metric=mesh.fortran_arrays(),
                   state=state, connections=list(mesh.connections))


def make_forcing(prob, seed=0, ft_scale=1e-4, fm_scale=1e-5, fq_scale=1e-7):
    """Seeded smooth CAM-style forcing tendencies for ``prob`` (K/s, m/s^2,
    kg/m^2/s per level): a solid-body-rotation FM, an l<=2 pattern in FT and
    per-tracer patterns in FQ whose negative parts exercise the clip."""
    rng = np.random.default_rng(seed)
    m = prob.mesh
    x, y, z = m.pos[..., 0], m.pos[..., 1], m.pos[..., 2]
    lat = m.lat
    nl = prob.nlev
    kk = np.linspace(0.5, 1.0, nl)[None, :, None, None]
    FT = ft_scale * (z + 0.5 * x * y)[:, None] * kk
    FM = np.zeros((prob.nelem, nl, 2, 4, 4))
    FM[:, :, 0] = fm_scale * np.cos(lat)[:, None] * kk
    FM[:, :, 1] = 0.3 * fm_scale * (x * z)[:, None] * kk
    FQ = np.zeros((prob.nelem, prob.qsize_d, nl, 4, 4))
    for q in range(prob.qsize):
        c = rng.standard_normal(4)
        FQ[:, q] = fq_scale * (c[0] + c[1] * x + c[2] * y + c[3] * z)[:, None] * kk
    return {"FM": np.ascontiguousarray(FM), "FT": np.ascontiguousarray(FT), "FQ": np.ascontiguousarray(FQ)}


# init_diagnostics_c's arrays in argument order (Diagnostics.cpp:17-35):
# the Fortran elem%state%Q and elem%accum arrays, here with C shapes (last
# Fortran index first, np x np trailing)
DIAGNOSTIC_ARRAYS = ("q", "qvar", "qmass", "q1mass", "iener", "iener_wet", "kener", "pener")


def diagnostic_arrays(prob):
    """Zeroed diagnostic arrays for ``prob``: q (nelem, qsize_d, nlev, 4, 4),
    qvar and qmass (nelem, 4, qsize_d, 4, 4) with the slot index 0..3 second,
    q1mass (nelem, qsize_d, 4, 4), iener, kener, pener (nelem, 4, 4, 4) and
    iener_wet (nelem, 4, 4)."""
    E, L, qd = prob.nelem, prob.nlev, prob.qsize_d
    shapes = {"q": (E, qd, L, 4, 4), "qvar": (E, 4, qd, 4, 4), "qmass": (E, 4, qd, 4, 4), "q1mass": (E, qd, 4, 4),
              "iener": (E, 4, 4, 4), "iener_wet": (E, 4, 4), "kener": (E, 4, 4, 4), "pener": (E, 4, 4, 4)}
    return {k: np.zeros(shapes[k]) for k in DIAGNOSTIC_ARRAYS}
