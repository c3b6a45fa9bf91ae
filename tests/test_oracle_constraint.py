"""Pins for the oracle's bulk constraint treatment (oracle/channel.py
constraint_treatment, restating suzerain/treatment_constraint.cpp:87-316
with the bulk constraints of constraint.cpp) and for the bulk weights.

* integration weights: the exact integrals of test_bspline.c:108-330
  (orders 2, 3, 4), the closed form (t_{i+k} - t_i) / k and the partition
  of unity (sum = L), for the oracle's Gauss-Legendre restatement and the
  product's closed form;
* the treatment: after it the bulk of every enabled field equals its target;
  the correction of the (0,0) mode, pushed back through the (0,0) operator
  (the separately pinned matrix-free imexop), is the forcing the treatment
  promises at every interior collocation point -- uniform in the
  constrained equation, u Ma^2 times the momentum forcing added to the
  energy equation -- and satisfies the homogeneous isothermal wall rows;
  disabled constraints change nothing;
* a constrained SMR91 step holds the bulk targets, and the running means
  of the implicit forcing are the plain average of the three substeps'
  scaled multipliers (a planted wrong average weight is caught).
"""
import ctypes as C
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle import channel
from oracle.bspline import breakpoints, collocation, integration_coefficients, knots

HERE = os.path.dirname(os.path.abspath(__file__))
SCEN = {"Re": 300.0, "Pr": 0.7, "Ma": 1.5, "alpha": 0.5, "beta": 2.0 / 3.0,
        "gamma": 1.4}
E, MX, MY, MZ, RHO = range(5)


def _golden():
    with open(os.path.join(HERE, "golden", "bspline_integration_golden.json")) as f:
        return json.load(f)


def _product_weights(k, bp):
    from suzerain_amd._lib import load
    lib = load()
    bp = np.ascontiguousarray(bp, dtype=np.float64)
    info = C.c_int(0)
    lib.sz_bsplineop_alloc.restype = C.c_void_p
    lib.sz_bsplineop_alloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int,
                                       C.c_void_p]
    w = lib.sz_bsplineop_alloc(k, len(bp), bp.ctypes.data, 0, C.byref(info))
    assert w, info.value
    n = len(bp) + k - 2
    out = np.empty(n)
    f = lib.sz_bsplineop_integration_coefficients
    f.argtypes = [C.c_void_p, C.c_void_p]
    assert f(w, out.ctypes.data) == 0
    lib.sz_bsplineop_free.argtypes = [C.c_void_p]
    lib.sz_bsplineop_free(w)
    return out


@pytest.mark.parametrize("case", _golden()["cases"], ids=lambda c: f"k{c['k']}")
def test_integration_weights_golden(case):
    gold = _golden()
    bp = [float(Fraction(b)) for b in gold["breakpoints"]]
    want = np.array([float(Fraction(v)) for v in case["e0"]])
    tol = gold["tolerance_eps_multiple"] * np.finfo(float).eps
    for got in (integration_coefficients(case["k"], bp),
                _product_weights(case["k"], bp)):
        np.testing.assert_allclose(got, want, rtol=tol, atol=0)


@pytest.mark.parametrize("k,ndof,htdelta", [(8, 96, 3.0), (6, 48, 2.0),
                                            (8, 192, 3.0)])
def test_integration_weights_channel_grid(k, ndof, htdelta):
    bp = breakpoints(ndof, k, 0.0, 2.0, htdelta)
    quad = integration_coefficients(k, list(bp))
    t = knots(k, bp)
    closed = (t[k:] - t[:-k]) / k
    prod = _product_weights(k, bp)
    assert abs(quad.sum() - 2.0) < 1e-14
    # quadrature rounding is absolute, a few eps of the largest weight
    tol = 1e-13 * closed.max()
    np.testing.assert_allclose(quad, closed, rtol=0, atol=tol)
    np.testing.assert_allclose(prod, quad, rtol=0, atol=tol)
    np.testing.assert_allclose(prod, closed, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- treatment
@pytest.fixture(scope="module")
def setup():
    k, ndof = 6, 14
    bp = breakpoints(ndof, k, -1.0, 1.0, 2.0)
    ops = oracle.Ops(collocation(k, list(bp), 2))
    g = channel.Grid(8, 8, 12, 12, 4 * math.pi, 4 * math.pi / 3, ops)
    weights = integration_coefficients(k, list(bp)) / 2.0
    from test_oracle_channel import random_state
    W = random_state(g)
    _, refs, _ = channel.apply_ns(g, W, SCEN, "rhome_xyz")
    R = oracle.ns.imexop_refs(refs)
    return g, weights, W, refs, R


def _scen5():
    return [SCEN[q] for q in ("Re", "Pr", "Ma", "alpha", "gamma")]


BC = ((1.0, 0, 0, 0), (1.0, 0, 0, 0))


def _solved(setup, phi):
    g, weights, W, refs, R = setup
    rng = np.random.default_rng(11)
    x = W * 0
    x[..., g.keep, :] = (rng.standard_normal(x[..., g.keep, :].shape) + 0.1) * 0.2
    x[:, 0, 0, :] = x[:, 0, 0, :].real + np.array([3, 0.8, 0.01, 0.05, 1.2])[:, None]
    return channel.invert_operator(g, phi, _scen5(), R, x, BC)


def _uvw(g):
    y = np.linspace(-1, 1, g.Ny)
    return np.array([1.2 * (1 - y * y), 0.03 * y, 0.05 * (1 - y * y)])


TARGETS = [2.9, 0.7, math.nan, math.nan, 1.15]


def test_treatment_meets_targets(setup):
    g, weights, W, refs, R = setup
    phi = -2e-3 * 111 / 480
    out = _solved(setup, phi)
    before = out.copy()
    cphi = channel.constraint_treatment(g, phi, _scen5(), R, BC, out, TARGETS,
                                        weights, _uvw(g), SCEN["Ma"])
    print("multipliers", cphi)
    for f, t in enumerate(TARGETS):
        bulk = np.dot(weights, out[f, 0, 0].real)
        if t == t:
            assert abs(bulk - t) <= 1e-13 * abs(t), (f, bulk, t)
    assert (cphi[[MY, MZ]] == 0).all() and (cphi[5:] == 0).all()
    assert (cphi[[E, MX, RHO]] != 0).all()
    # only the (0,0) mode moves
    diff = out - before
    diff[:, 0, 0] = 0
    assert not diff.any()


def test_correction_is_operator_response_to_forcing(setup):
    """(M + phi L) applied to the correction is the promised forcing in the
    interior; the wall rows hold the homogeneous isothermal constraints."""
    g, weights, W, refs, R = setup
    phi = -2e-3 * 100 / 480
    out = _solved(setup, phi)
    mean0 = out[:, 0, 0].copy()
    uvw = _uvw(g)
    Ma = SCEN["Ma"]
    cphi = channel.constraint_treatment(g, phi, _scen5(), R, BC, out, TARGETS,
                                        weights, uvw, Ma)
    delta = np.ascontiguousarray(out[:, 0, 0] - mean0)
    y = np.zeros_like(delta)
    oracle.imexop_accumulate(phi, 0.0, 0.0, _scen5(), R, g.ops, delta, 0.0, y)
    want = np.zeros((5, g.Ny))
    want[E] = cphi[E] + (cphi[MX] * uvw[0] + cphi[MY] * uvw[1]
                         + cphi[MZ] * uvw[2]) * Ma * Ma
    want[MX], want[MY], want[MZ], want[RHO] = cphi[MX], cphi[MY], cphi[MZ], cphi[RHO]
    inner = slice(1, g.Ny - 1)
    scale = np.abs(want).max()
    assert np.abs(y[:4, inner] - want[:4, inner]).max() <= 1e-12 * scale
    assert np.abs(y[RHO] - want[RHO]).max() <= 1e-12 * scale
    # isothermal walls: rho E = rho / (gamma (gamma - 1)) at T = 1 and no
    # momentum, for the homogeneous (zeroed) wall rows
    gamma = SCEN["gamma"]
    Ef = 1.0 / (gamma * (gamma - 1))
    dscale = np.abs(delta).max()
    for j in (0, g.Ny - 1):
        assert abs(delta[E, j] - Ef * delta[RHO, j]) <= 1e-12 * dscale
        assert np.abs(delta[1:4, j]).max() <= 1e-12 * dscale


def test_disabled_constraints_change_nothing(setup):
    g, weights, W, refs, R = setup
    phi = -1e-3
    out = _solved(setup, phi)
    before = out.copy()
    cphi = channel.constraint_treatment(g, phi, _scen5(), R, BC, out,
                                        [math.nan] * 5, weights, _uvw(g), SCEN["Ma"])
    assert not cphi.any()
    np.testing.assert_array_equal(out, before)


def _lam(g):
    y = np.linspace(-1, 1, g.Ny)
    return channel.lambdas(g, y, 6, 2.0, 1.0, 1.0)


def test_constrained_step_holds_bulk_and_averages_forcing(setup):
    g, weights, W, refs, R = setup
    L = channel.HybridOperator(g, _scen5(), None, BC, bulk=TARGETS,
                               weights=weights, Ma=SCEN["Ma"])
    a, dt, _ = channel.smr91_step(g, W, SCEN, BC, _lam(g), max_dt=1e-3, L=L)
    assert dt == 1e-3
    for f, t in enumerate(TARGETS):
        if t == t:
            assert abs(np.dot(weights, a[f, 0, 0].real) - t) <= 1e-13 * abs(t)
    free, _, _ = channel.smr91_step(g, W, SCEN, BC, _lam(g), max_dt=1e-3)
    drift = [abs(np.dot(weights, free[f, 0, 0].real) - TARGETS[f]) for f in (E, MX, RHO)]
    assert min(drift) > 1e-6, drift                  # unconstrained: off target
    assert len(L.history) == 3
    scale = 3 / dt
    samples = [(c * scale, uvw) for c, uvw in L.history]
    fx = np.mean([c[MX] for c, _ in samples])
    qb = np.mean([c[E] for c, _ in samples])
    crho = np.mean([c[RHO] for c, _ in samples])
    c2rho = np.mean([c[RHO] ** 2 for c, _ in samples])
    fdu = np.mean([c[MX] * u[0] + c[MY] * u[1] + c[MZ] * u[2]
                   for c, u in samples], axis=0)
    for name, want in (("fx", fx), ("qb", qb), ("Crho", crho), ("C2rho", c2rho)):
        np.testing.assert_allclose(L.imp[name], want, rtol=1e-13, atol=0)
    np.testing.assert_allclose(L.imp["f_dot_u"], fdu, rtol=1e-12, atol=1e-300)
    for name in ("fy", "fz", "CrhoE", "C2rhoE", "Crhou", "Crhou_dot_u"):
        assert not L.imp[name].any(), name
    # a running mean with weight 1/(i+2) instead of 1/(i+1) misses the plain
    # average: the check above would catch it
    wrong = 0.0
    for i, (c, _) in enumerate(samples):
        wrong += (c[MX] - wrong) / (i + 2)
    assert abs(wrong - fx) > 1e-6 * abs(fx)
