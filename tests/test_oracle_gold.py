"""The reference's own regression states pin the oracle's whole SMR91 step.

apps/perfect/manual.gold/channel.rhome_xyz holds a channel (Ny = 32, k = 8,
htdelta = 2, Re = 100, Ma = 1.5, bulk rho and rho u driven to 1) before and
after perfect_advance --implicit=rhome_xyz advanced it to t = 0.000919
(check.sh).  The fixture tests/golden/channel_rhome_xyz_gold.json holds both
states (transcribed by tests/golden/make_gold_state.py).

* the oracle's breakpoints, collocation points and integration weights are
  the reference's;
* the reference's initial state is the isothermal laminar profile of
  laminar_state (parabolic u of unit bulk velocity, rho = T = 1);
* one oracle SMR91 step with the bulk constraint treatment reproduces the
  reference's advanced state, per field <= 1e-12 of the field's magnitude
  (the wall-normal momentum, six orders of magnitude below the others and
  born of rounding-sized density changes, <= 1e-10 of its own magnitude);
  without the constraint treatment the step misses it.

The reference's saved running means of the implicit forcing (bar_f ...) are
not compared: they come out of the driver's statistics bookkeeping, and the
per-substep average of the treatment (treatment_constraint.cpp:226-316)
does not reproduce them from this single step.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from oracle import channel as och
from oracle.bspline import breakpoints, collocation, integration_coefficients

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("rho_E", "rho_u", "rho_v", "rho_w", "rho")


def load_gold():
    with open(os.path.join(HERE, "golden", "channel_rhome_xyz_gold.json")) as fh:
        return json.load(fh)


def _f(v):
    return math.nan if v == "nan" else float(v)


def gold_state(case):
    """(5, Ny) complex B-spline coefficients of the (0,0) mode."""
    out = []
    for name in FIELDS:
        s = case["state"][name]
        assert s["shape"][:2] == [1, 1]
        out.append(np.array(s["re"]) + 1j * np.array(s["im"]))
    return np.array(out)


def gold_setup(gold, dN=2):
    ini = gold["initial"]
    k, Ny = ini["k"], ini["Ny"]
    bp = breakpoints(Ny, k, 0.0, ini["Ly"], ini["htdelta"])
    coll = collocation(k, list(bp), 2)
    ops = oracle.Ops(coll)
    g = och.Grid(ini["Nx"], ini["Nz"], dN, dN, ini["Lx"], ini["Lz"], ops)
    sc = {q: ini[q] for q in ("Re", "Pr", "Ma", "alpha", "beta", "gamma")}
    bc = tuple(tuple(ini[f"{w}_{q}"] for q in ("T", "u", "v", "w"))
               for w in ("lower", "upper"))
    targets = [_f(ini["bulk_rho_E"]), _f(ini["bulk_rho_u"]), math.nan, math.nan,
               _f(ini["bulk_rho"])]
    return g, coll, bp, sc, bc, targets


def field_errors(got, want):
    return [float(np.abs(got[f] - want[f]).max() / max(np.abs(want[f]).max(), 1e-300))
            for f in range(5)]


# per-field bars: the wall-normal momentum is rounding-born (see above)
BARS = (1e-12, 1e-12, 1e-10, 0.0, 1e-12)


def test_grid_matches_reference():
    gold = load_gold()
    ini = gold["initial"]
    g, coll, bp, *_ = gold_setup(gold)
    np.testing.assert_array_equal(bp, np.array(ini["breakpoints_y"]))
    np.testing.assert_allclose(coll["points"], ini["collocation_points_y"],
                               rtol=0, atol=1e-15)
    w = integration_coefficients(ini["k"], list(bp))
    np.testing.assert_allclose(w, ini["integration_weights"], rtol=0,
                               atol=1e-15 * w.max())


def test_initial_state_is_laminar_profile():
    gold = load_gold()
    ini = gold["initial"]
    g, coll, *_ = gold_setup(gold)
    M, y = coll["D"][0], coll["points"]
    vals = M @ gold_state(ini).real.T                     # (Ny, 5)
    rho, mx, e = vals[:, 4], vals[:, 1], vals[:, 0]
    u = mx / rho
    Ly = ini["Ly"]
    gamma, Ma = ini["gamma"], ini["Ma"]
    assert np.abs(rho - 1).max() < 1e-14
    assert np.abs(u - 6 / Ly ** 2 * y * (Ly - y)).max() < 1e-14
    assert np.abs(e - rho * (1 / (gamma * (gamma - 1)) + Ma * Ma / 2 * u * u)).max() < 1e-14


@pytest.mark.parametrize("dN", [1, 2])
def test_constrained_step_reproduces_reference(dN):
    gold = load_gold()
    ini, rst = gold["initial"], gold["restart"]
    g, coll, bp, sc, bc, targets = gold_setup(gold, dN)
    weights = integration_coefficients(ini["k"], list(bp)) / ini["Ly"]
    W0 = np.zeros((5, g.dNz, g.Xw, g.Ny), dtype=complex)
    W0[:, 0, 0] = gold_state(ini)
    want = gold_state(rst)
    scen5 = [sc[q] for q in ("Re", "Pr", "Ma", "alpha", "gamma")]
    lam = och.lambdas(g, coll["points"], ini["k"], ini["htdelta"], 1.0, 1.0)
    L = och.HybridOperator(g, scen5, None, bc, bulk=targets, weights=weights,
                           Ma=sc["Ma"])
    a, dt, _ = och.smr91_step(g, W0, sc, bc, lam, max_dt=rst["t"], L=L)
    assert dt == rst["t"]
    errs = field_errors(a[:, 0, 0], want)
    print(f"dN={dN}: oracle step vs reference per field {errs}")
    for f in range(5):
        assert errs[f] <= BARS[f], (f, errs)
    # the unconstrained step misses the reference state
    free, _, _ = och.smr91_step(g, W0, sc, bc, lam, max_dt=rst["t"])
    ferr = field_errors(free[:, 0, 0], want)
    assert ferr[1] > 1e-9, ferr
