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

    @property
    def nelem(self):
        return self.mesh.nelem

    @property
    def qsize_d(self):
        return self.state["qsize_d"]


_MESH_CACHE = {}


def build_problem(ne=4, nlev=26, qsize=4, seed=0, dt=300.0, t_noise=0.0, u_perturb=1.0, **param_overrides):
    """Synthetic JW-baroclinic problem on an ``ne`` cubed sphere.

    Defaults follow the small CPU configuration (ne=4, L26, Q4, dt=300 s,
    nu=1e16, hypervis subcycle 3, remap every step, limiter 8, remap alg 1,
    no forcing).  ``u_perturb`` scales the 1 m/s Gaussian jet perturbation
    (0 gives the steady balanced jet).
    """
    if ne not in _MESH_CACHE:
        _MESH_CACHE[ne] = CubedSphere(ne)
    mesh = _MESH_CACHE[ne]
    vc = VCoord(nlev)
    p = Params(qsize=qsize, **param_overrides)
    state = jw_baroclinic(mesh, vc, qsize, seed=seed, t_noise=t_noise, u_perturb=u_perturb)
    return Problem(ne=ne, nlev=nlev, qsize=qsize, dt=dt, params=p, mesh=mesh, vcoord=vc,
                   dvv=np.ascontiguousarray(derivative_matrix()), metric=mesh.fortran_arrays(),
                   state=state, connections=list(mesh.connections))
