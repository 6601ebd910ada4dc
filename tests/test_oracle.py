"""Pins the CPU oracle before it is trusted as the parity reference.

* sphere operators: the reference's own golden vectors
  (test/unit_tests/inputs/*_sphere_np4.in, copied under tests/golden/), bit-exact;
* sphere operators on the real mesh: closed forms on the sphere (gradient of
  sin(lat), vorticity of solid-body rotation, the Y_1 eigenvalue of the
  Laplacian) and the exact discrete identity sum(weak laplacian) == 0;
* DSS: assembled fields are continuous, DSS(spheremp*c)*rspheremp == c, and
  the unique-node count of an np=4 cube mesh is 54 ne^2 + 2;
* limiters 8 and 9: the property checks of the reference's
  preqx_ut.cpp:1335-1532 (bounds, mass, equal 1-norm of the two corrections);
* PPM remap: identity remap, column-mass conservation, constants preserved,
  monotone data stays within its range;
* full step: dry-mass and tracer-mass conservation and an isothermal resting
  atmosphere that stays at rest.
"""
import math
import os

import numpy as np
import pytest

import oracle
from hommexx_amd.problem import build_problem

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_inputs")
EPS = np.finfo(float).eps
A = 6.376e6


def parse_in(name):
    d, key = {}, None
    for ln in open(os.path.join(GOLD, name)):
        t = ln.split()
        if not t:
            continue
        try:
            vals = [float(x) for x in t]
        except ValueError:
            key = ln.strip()
            continue
        d.setdefault(key, []).extend(vals)
    return {k: np.array(v) for k, v in d.items()}


P = oracle.ptr


def test_golden_gradient(oracle_lib):
    g = parse_in("gradient_sphere_np4.in")
    out = np.zeros(32)
    oracle_lib.oracle_unit_gradient(1, P(g["deriv_Dvv"]), P(g["elem_Dinv"]), P(g["s"]), P(out))
    assert np.array_equal(out, g["Gradient Sphere result"])


def test_golden_divergence(oracle_lib):
    d = parse_in("divergence_sphere_np4.in")
    out = np.zeros(16)
    oracle_lib.oracle_unit_divergence(1, P(d["deriv_Dvv"]), P(d["elem_Dinv"]), P(d["elem_metdet"]), P(d["v"]),
                                      P(out))
    assert np.array_equal(out, d["Divergence Sphere result"])


def test_golden_vorticity(oracle_lib):
    v = parse_in("vorticity_sphere_np4.in")
    out = np.zeros(16)
    md = 1.0 / v["elem_rmetdet"]
    oracle_lib.oracle_unit_vorticity(1, P(v["deriv_Dvv"]), P(v["elem_D"]), P(md), P(v["v"]), P(out))
    assert np.array_equal(out, v["Vorticity Sphere result"])


# ------------------------------------------------------------- mesh operators
@pytest.fixture(scope="module")
def mesh_prob():
    return build_problem(ne=4, nlev=26, qsize=2, seed=3)


def _elem(arr, ie):
    return np.ascontiguousarray(arr[ie], dtype=np.float64)


def test_gradient_of_sin_lat(oracle_lib, mesh_prob):
    m, pr = mesh_prob.metric, mesh_prob
    lat = pr.mesh.lat
    err = 0.0
    for ie in range(pr.nelem):
        s = np.ascontiguousarray(np.sin(lat[ie]))
        g = np.zeros(32)
        oracle_lib.oracle_unit_gradient(1, P(pr.dvv), P(_elem(m["Dinv"], ie)), P(s), P(g))
        exact_lat = np.cos(lat[ie]) / A
        err = max(err, np.abs(g[:16].reshape(4, 4)).max() * A, np.abs(g[16:].reshape(4, 4) - exact_lat).max() * A)
    assert err < 2e-3  # spectral element truncation error at ne=4


def _rotation_errors(oracle_lib, pr):
    m = pr.metric
    lat = pr.mesh.lat
    ev = ed = 0.0
    for ie in range(pr.nelem):
        v = np.ascontiguousarray(np.concatenate([np.cos(lat[ie]).ravel(), np.zeros(16)]))
        vort = np.zeros(16)
        oracle_lib.oracle_unit_vorticity(1, P(pr.dvv), P(_elem(m["D"], ie)), P(_elem(m["metdet"], ie)), P(v),
                                         P(vort))
        ev = max(ev, np.abs(vort * A - 2.0 * np.sin(lat[ie]).ravel()).max())
        div = np.zeros(16)
        oracle_lib.oracle_unit_divergence(1, P(pr.dvv), P(_elem(m["Dinv"], ie)), P(_elem(m["metdet"], ie)), P(v),
                                          P(div))
        ed = max(ed, np.abs(div * A).max())
    return ev, ed


def test_vorticity_of_solid_body_rotation(oracle_lib, mesh_prob):
    """curl of u = cos(lat) (zonal) is 2 sin(lat)/a and its divergence is 0;
    element-local errors shrink at third order as the mesh is refined."""
    e4 = _rotation_errors(oracle_lib, mesh_prob)
    e8 = _rotation_errors(oracle_lib, build_problem(ne=8, nlev=26, qsize=1))
    assert max(e4) < 2e-2
    assert e4[0] / e8[0] > 5 and e4[1] / e8[1] > 5


def _weak_laplace_all(oracle_lib, pr, f):
    m = pr.metric
    out = np.zeros_like(f)
    for ie in range(pr.nelem):
        lap = np.zeros(16)
        oracle_lib.oracle_unit_laplace(1, P(pr.dvv), P(_elem(m["Dinv"], ie)), P(_elem(m["spheremp"], ie)),
                                       P(np.ascontiguousarray(f[ie])), P(lap))
        out[ie] = lap.reshape(4, 4)
    return out


def test_weak_laplacian_sums_to_zero_and_y1_eigenvalue(oracle_lib, mesh_prob):
    pr = mesh_prob
    rng = np.random.default_rng(0)
    f = rng.standard_normal((pr.nelem, 4, 4))
    lap = _weak_laplace_all(oracle_lib, pr, f)
    assert abs(lap.sum()) <= 1e-12 * np.abs(lap).sum()
    # Y_1 = sin(lat): DSS'd weak laplacian / mass = -2/a^2 sin(lat)
    def y1_error(p):
        z = np.sin(p.mesh.lat)
        lapz = _weak_laplace_all(oracle_lib, p, z)
        assembled = p.mesh.dss_scalar_reference(lapz) * p.mesh.rspheremp
        return np.abs(assembled * A * A + 2.0 * z).max() / 2.0

    e4 = y1_error(pr)
    e8 = y1_error(build_problem(ne=8, nlev=26, qsize=1))
    assert e4 < 3e-2 and e4 / e8 > 3


# ----------------------------------------------------------------------- DSS
def test_dss_continuity_and_weighted_constant(oracle_lib, mesh_prob):
    pr = mesh_prob
    s = oracle.OracleSession(pr)
    nl = pr.nlev
    rng = np.random.default_rng(5)
    f = rng.standard_normal((2, pr.nelem, nl, 4, 4))
    # field 1: spheremp * constant per level
    c = rng.uniform(1, 2, size=nl)
    f[1] = pr.metric["spheremp"][:, None, :, :] * c[None, :, None, None]
    oracle_lib.oracle_unit_dss(s.h, 2, nl, P(f), 1)
    assert np.abs(f[1] / c[None, :, None, None] - 1).max() < 8 * EPS
    # continuity: all copies of a node agree after DSS (geometric matching)
    key = np.round(pr.mesh.pos * 1e9).astype(np.int64).reshape(-1, 3)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    assert inv.max() + 1 == 54 * pr.ne ** 2 + 2
    vals = np.transpose(f[0], (1, 0, 2, 3)).reshape(nl, -1)
    for k in range(0, nl, 7):
        lo = np.full(inv.max() + 1, np.inf)
        hi = np.full(inv.max() + 1, -np.inf)
        np.minimum.at(lo, inv, vals[k])
        np.maximum.at(hi, inv, vals[k])
        assert np.max(hi - lo) <= 1e-13 * np.abs(vals[k]).max()
    s.close()


# ------------------------------------------------------------------- limiters
def _feasible(rng, nlev):
    sph = rng.uniform(1.0 / 16, 2.0 / 16, 16)
    dpm = rng.uniform(0.5, 1.0, (16, nlev))
    pt = rng.uniform(0, 1, (16, nlev))
    ql = np.sort(rng.uniform(0, 1, (2, nlev)), axis=0)
    pt *= dpm
    for k in range(nlev):
        m = np.sum(sph * pt[:, k])
        lo = np.sum(sph * ql[0, k] * dpm[:, k])
        hi = np.sum(sph * ql[1, k] * dpm[:, k])
        dm = lo - m if m < lo else (hi - m if m > hi else 0)
        if dm:
            pt[:, k] += (1 + 1e2 * EPS) * dm / (16 * sph)
    mass = np.einsum("p,pk->k", sph, pt)
    return sph, np.ascontiguousarray(dpm), np.ascontiguousarray(ql.copy()), np.ascontiguousarray(pt), mass


@pytest.mark.parametrize("option", [8, 9])
def test_limiter_bounds_and_mass(oracle_lib, option):
    rng = np.random.default_rng(11 + option)
    nlev = 64
    sph, dpm, ql, pt, mass = _feasible(rng, nlev)
    oracle_lib.oracle_unit_limiter(option, nlev, P(sph), P(dpm), P(ql), P(pt))
    assert np.all(pt >= (1 - 1e1 * EPS) * ql[0] * dpm)
    assert np.all(pt <= (1 + 1e1 * EPS) * ql[1] * dpm)
    m = np.einsum("p,pk->k", sph, pt)
    assert np.all(np.abs(m - mass) <= 1e2 * EPS * mass)


def test_limiters_have_equal_one_norm(oracle_lib):
    rng = np.random.default_rng(7)
    nlev = 64
    sph, dpm, ql, pt, mass = _feasible(rng, nlev)
    a, b = pt.copy(), pt.copy()
    oracle_lib.oracle_unit_limiter(8, nlev, P(sph), P(dpm), P(ql.copy()), P(a))
    oracle_lib.oracle_unit_limiter(9, nlev, P(sph), P(dpm), P(ql.copy()), P(b))
    n8 = np.einsum("p,pk->k", sph, np.abs(a - pt))
    n9 = np.einsum("p,pk->k", sph, np.abs(b - pt))
    assert np.all(np.abs(n8 - n9) <= 1e3 * EPS * np.maximum(n8, 1e-300))


# ------------------------------------------------------------------------ PPM
def _ppm(oracle_lib, src, tgt, vals, alg=1):
    v = np.ascontiguousarray(vals.copy())
    oracle_lib.oracle_unit_ppm(len(src), alg, P(np.ascontiguousarray(src)), P(np.ascontiguousarray(tgt)),
                               v.shape[0], P(v))
    return v


@pytest.mark.parametrize("alg", [1, 2, 3])
def test_ppm_identity_constant_and_mass(oracle_lib, alg):
    rng = np.random.default_rng(alg)
    n = 72
    src = rng.uniform(100, 1000, n)
    q = rng.uniform(0.1, 1.0, (3, n)) * src
    out = _ppm(oracle_lib, src, src, q, alg)
    assert np.abs(out - q).max() <= 1e-12 * np.abs(q).max()
    tgt = src * rng.uniform(0.9, 1.1, n)
    tgt *= src.sum() / tgt.sum()
    const = np.vstack([src * 0.7])
    oc = _ppm(oracle_lib, src, tgt, const, alg)
    assert np.abs(oc / tgt - 0.7).max() < 1e-12
    out = _ppm(oracle_lib, src, tgt, q, alg)
    assert np.abs(out.sum(axis=1) - q.sum(axis=1)).max() <= 1e-12 * q.sum(axis=1).max()


def test_ppm_monotone_stays_in_range(oracle_lib):
    rng = np.random.default_rng(9)
    n = 72
    src = rng.uniform(100, 1000, n)
    tgt = src * rng.uniform(0.8, 1.2, n)
    tgt *= src.sum() / tgt.sum()
    mix = np.cumsum(rng.uniform(0, 1, n))
    mix /= mix[-1]
    out = _ppm(oracle_lib, src, tgt, (mix * src)[None, :], 1)[0] / tgt
    assert out.min() >= mix.min() - 1e-12 and out.max() <= mix.max() + 1e-12


# ------------------------------------------------------------------ full step
def test_full_step_conserves_mass(oracle_lib):
    pr = build_problem(ne=4, nlev=26, qsize=3, seed=2)
    s = oracle.OracleSession(pr)
    sph = pr.metric["spheremp"]
    st0 = s.state()
    dry0 = (sph * st0["ps_v"][:, 1]).sum()
    q0 = np.einsum("epq,ecklpq->c", sph, st0["Qdp"][:, :1])[: pr.qsize]
    for _ in range(3):
        nstep, nm1, n0, np1 = s.run(1)
    st = s.state()
    dry = (sph * st["ps_v"][:, n0]).sum()
    q = np.einsum("epq,ecklpq->c", sph, st["Qdp"][:, nstep % 2][:, None])[: pr.qsize]
    assert abs(dry / dry0 - 1) < 1e-14
    assert np.all(np.abs(q / q0 - 1) < 1e-14)
    s.close()


def test_isothermal_rest_state_stays_at_rest(oracle_lib):
    pr = build_problem(ne=4, nlev=26, qsize=1, seed=0)
    st = pr.state
    st["v"][:] = 0.0
    st["T"][:] = 250.0
    st["phis"][:] = 0.0
    Q = 0.3
    st["Qdp"][:] = Q * st["dp3d"][:, :2, None]
    s = oracle.OracleSession(pr)
    nstep, nm1, n0, np1 = s.run(2)
    out = s.state()
    assert np.abs(out["v"][:, n0]).max() < 1e-9
    assert np.abs(out["T"][:, n0] - 250.0).max() < 1e-9
    assert np.abs(out["ps_v"][:, n0] / 1e5 - 1).max() < 1e-13
    assert np.abs(out["Q"][:, 0] - Q).max() < 1e-13
    s.close()


def test_synthetic_vertical_coordinate():
    """Level counts without a table get the 72-level interfaces interpolated
    in index: same top and bottom, strictly increasing pressure at p0,
    midpoints between their interfaces."""
    from hommexx_amd.ics import VCoord
    ref = VCoord(72)
    for nlev in (3, 100, 160, 256):
        vc = VCoord(nlev)
        assert len(vc.hyai) == nlev + 1 and len(vc.hyam) == nlev
        assert vc.hyai[0] == ref.hyai[0] and vc.hybi[-1] == ref.hybi[-1] and vc.hybi[0] == ref.hybi[0]
        pi = vc.hyai * vc.ps0 + vc.hybi * vc.ps0
        assert np.all(np.diff(pi) > 0)
        pm = vc.hyam * vc.ps0 + vc.hybm * vc.ps0
        assert np.all((pm > pi[:-1]) & (pm < pi[1:]))
