"""Synthetic initial states for the preqx step (host-side set-up).

``jw_baroclinic`` restates the Jablonowski-Williamson balanced jet of the
reference (src/test_src/baroclinic_inst_mod.F90:56-271): zonal jet
``u = u0 cos^{3/2}(eta_v) sin^2(2 lat)`` plus a 1 m/s Gaussian bump of radius
1/10 (radians) centred at (20E, 40N), temperature in gradient-wind balance and
the matching surface geopotential, ``ps = p0``.

Tracers are seeded smooth fields in [0, 1] built from low-order (l <= 2) real
spherical harmonics, modulated in the vertical.  ``Qdp = Q * dp3d`` with
``dp3d = dhyai*ps0 + dhybi*ps``.

All arrays are returned in the flat Fortran layouts the C-ABI consumes
(``init_elements_states_c``): numpy C-order shapes equal to the reversed
Fortran shapes, e.g. ``v`` is ``(nelem, 3, nlev, 2, 4, 4)`` for Fortran
``elem_state_v(np,np,2,nlev,3,nelemd)``.
"""
import json
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

RGAS = 287.04
G = 9.80616
OMEGA = 7.292e-5
REARTH = 6.376e6
P0 = 100000.0


class VCoord:
    """Hybrid sigma-pressure coefficients (interfaces ``hyai/hybi``, midpoints
    ``hyam/hybm``), read from the package's JSON tables (L26, L72).  Other
    level counts (at least 3) get a synthetic coordinate: the 72-level
    interfaces interpolated linearly in the interface index onto nlev + 1
    points, midpoints as interface averages."""

    def __init__(self, nlev, ps0=P0):
        data = os.path.join(os.path.dirname(__file__), "data")
        path = os.path.join(data, "vcoord_L%d.json" % nlev)
        self.nlev = nlev
        self.ps0 = float(ps0)
        if os.path.exists(path):
            d = json.load(open(path))
            self.hyai = np.array(d["hyai"], float)
            self.hybi = np.array(d["hybi"], float)
            self.hyam = np.array(d["hyam"], float)
            self.hybm = np.array(d["hybm"], float)
            return
        if nlev < 3:
            raise ValueError("no vertical coordinate for nlev=%d" % nlev)
        d = json.load(open(os.path.join(data, "vcoord_L72.json")))
        x = np.linspace(0.0, 72.0, nlev + 1)
        self.hyai = np.interp(x, np.arange(73.0), np.array(d["hyai"], float))
        self.hybi = np.interp(x, np.arange(73.0), np.array(d["hybi"], float))
        self.hyam = 0.5 * (self.hyai[1:] + self.hyai[:-1])
        self.hybm = 0.5 * (self.hybi[1:] + self.hybi[:-1])

    @property
    def etam(self):
        return self.hyam + self.hybm

    def dp3d(self, ps):
        """Layer thickness from surface pressure, in the reference's order
        ``dhyai*ps0 + dhybi*ps`` (prim_driver.cpp:101-110)."""
        da = (self.hyai[1:] - self.hyai[:-1]).reshape(-1, 1, 1)
        db = (self.hybi[1:] - self.hybi[:-1]).reshape(-1, 1, 1)
        ps = np.asarray(ps)[..., None, :, :]
        return da * self.ps0 + db * ps


def _by_elements(fn, nelem):
    """fn(lo, hi) over element ranges on a few host threads (numpy releases
    the GIL in its array loops); for the multi-GB tracer arrays of large
    meshes"""
    nt = max(1, min(16, os.cpu_count() or 1, nelem // 256))
    if nt == 1:
        fn(0, nelem)
        return
    cuts = [nelem * i // nt for i in range(nt + 1)]
    with ThreadPoolExecutor(nt) as ex:
        list(ex.map(lambda i: fn(cuts[i], cuts[i + 1]), range(nt)))


def _smooth_tracers(mesh, nlev, qsize, seed, out=None):
    """Seeded smooth tracers in [0, 1]: a random l<=2 harmonic pattern per
    tracer, scaled into [0.05, 0.95], times a gentle vertical modulation."""
    rng = np.random.default_rng(seed)
    x, y, z = mesh.pos[..., 0], mesh.pos[..., 1], mesh.pos[..., 2]
    basis = [np.ones_like(x), x, y, z, x * y, y * z, x * z, x * x - y * y, 3 * z * z - 1]
    if out is None:
        out = np.empty((mesh.nelem, qsize, nlev, 4, 4))
    kk = np.arange(nlev)
    for q in range(qsize):
        c = rng.standard_normal(len(basis))
        c[0] = 0.0
        h = sum(ci * b for ci, b in zip(c, basis))
        hmax = np.abs(h).max()
        h = h / (hmax if hmax > 0 else 1.0)  # in [-1, 1]
        phase = rng.uniform(0, 2 * math.pi)
        vert = 0.8 + 0.2 * np.cos(2 * math.pi * kk / max(nlev, 1) + phase)  # in [0.6, 1]
        base = 0.5 + 0.45 * h  # in [0.05, 0.95]

        def fill(lo, hi, q=q, base=base, vert=vert):
            np.multiply(base[lo:hi, None, :, :], vert[None, :, None, None], out=out[lo:hi, q])

        _by_elements(fill, mesh.nelem)
    return out


def jw_baroclinic(mesh, vc, qsize, seed=0, qsize_d=None, t_noise=0.0, u_perturb=1.0):
    """Balanced JW jet state on ``mesh`` with ``vc`` levels and ``qsize`` tracers.

    ``t_noise`` (K) adds seeded white noise to T (used by the mixed-precision
    configuration).  Returns a dict of Fortran-layout arrays.
    """
    if qsize_d is None:
        qsize_d = max(qsize, 1)
    nlev = vc.nlev
    nelem = mesh.nelem
    u0, t0, gama, ddt = 35.0, 288.0, 0.005, 4.8e5
    eta_t, eta_s, eta_0 = 0.2, 1.0, 0.252
    pi = math.pi
    eta = vc.etam
    etv = (eta - eta_0) * pi * 0.5
    latc, lonc = pi * (2.0 / 9.0), pi * (1.0 / 9.0)
    lat, lon = mesh.lat, mesh.lon  # (nelem, j, i)
    snlat, cslat = np.sin(lat), np.cos(lat)
    aa = math.sin(latc) * snlat + math.cos(latc) * cslat * np.cos(lon - lonc)
    rc = 10.0 * np.arccos(np.clip(aa, -1.0, 1.0))
    v = np.zeros((nelem, 3, nlev, 2, 4, 4))
    for k in range(nlev):
        v1 = u0 * math.cos(etv[k]) ** 1.5 * np.sin(2.0 * lat) ** 2 + u_perturb * np.exp(-rc * rc)
        v[:, :, k, 0] = v1[:, None]
    tbar = np.where(eta <= eta_t, t0 * eta ** (RGAS * gama / G) + ddt * (eta_t - eta) ** 5,
                    t0 * eta ** (RGAS * gama / G))
    T = np.zeros((nelem, 3, nlev, 4, 4))
    for k in range(nlev):
        trm1 = 0.75 * (eta[k] * pi * u0 / RGAS) * math.sin(etv[k]) * math.sqrt(math.cos(etv[k]))
        trm2 = -2.0 * snlat ** 6 * (cslat ** 2 + 1.0 / 3.0) + 10.0 / 63.0
        trm3 = 2.0 * u0 * math.cos(etv[k]) ** 1.5
        trm4 = (1.60 * cslat ** 3 * (snlat ** 2 + 2.0 / 3.0) - pi * 0.25) * REARTH * OMEGA
        T[:, :, k] = (tbar[k] + trm1 * (trm2 * trm3 + trm4))[:, None]
    if t_noise:
        rng = np.random.default_rng(seed + 7919)
        T += t_noise * rng.standard_normal(T.shape[0:1] + T.shape[2:])[:, None]
    trm1 = u0 * math.cos((eta_s - eta_0) * pi * 0.5) ** 1.5
    trm2 = -2.0 * snlat ** 6 * (cslat ** 2 + 1.0 / 3.0) + 10.0 / 63.0
    trm3 = (1.60 * cslat ** 3 * (snlat ** 2 + 2.0 / 3.0) - pi * 0.25) * REARTH * OMEGA
    phis = np.ascontiguousarray(trm1 * (trm2 * trm1 + trm3))
    ps_v = np.full((nelem, 3, 4, 4), P0)
    dp = vc.dp3d(ps_v[:, 0])  # (nelem, nlev, 4, 4)
    dp3d = np.repeat(dp[:, None], 3, axis=1)
    Q = np.zeros((nelem, qsize_d, nlev, 4, 4))
    if qsize > 0:
        _smooth_tracers(mesh, nlev, qsize, seed, out=Q[:, :qsize])
    Qdp = np.empty((nelem, 2, qsize_d, nlev, 4, 4))

    def qdp(lo, hi):
        np.multiply(Q[lo:hi], dp[lo:hi, None], out=Qdp[lo:hi, 0])
        Qdp[lo:hi, 1] = Qdp[lo:hi, 0]

    _by_elements(qdp, nelem)
    return {
        "v": np.ascontiguousarray(v), "T": np.ascontiguousarray(T),
        "dp3d": np.ascontiguousarray(dp3d), "ps_v": ps_v, "phis": phis,
        "Qdp": Qdp, "Q": Q, "qsize_d": qsize_d,
    }
