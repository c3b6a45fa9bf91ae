"""Pins for the oracle's explicit channel operator (oracle/channel.py) and,
through it, the velocity-dependent terms of the implicit operator table.

The rhome_xyz right hand side is the fully explicit one minus the implicit
operator L built from the plane-mean references (navier_stokes.hpp:475-752
against rholut_imexop.c:43-531), so for ANY state
    N_none(u) - N_rhome_xyz(u) = (dNx dNz) L u
on every retained wavenumber, L applied to the B-spline coefficients and the
factor being the unnormalized forward transform.  The left side comes from
the pointwise formulas pinned to test_rholut.cpp's Sage values; the right
side from the 96-term operator table; the two are written independently.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import channel
from oracle.bspline import breakpoints, collocation

SCEN = {"Re": 300.0, "Pr": 0.7, "Ma": 1.5, "alpha": 0.5, "beta": 2.0 / 3.0,
        "gamma": 1.4}


@pytest.fixture(scope="module")
def grid():
    k, ndof = 6, 14
    bp = breakpoints(ndof, k, -1.0, 1.0, 2.0)
    ops = oracle.Ops(collocation(k, list(bp), 2))
    return channel.Grid(8, 8, 12, 12, 4 * math.pi, 4 * math.pi / 3, ops)


def random_state(g, seed=3, amp=0.004):
    """Positive, smooth state with every retained mode excited: mean
    profiles plus random perturbations; returned as wave coefficients."""
    rng = np.random.default_rng(seed)
    Ny = g.Ny
    y = np.linspace(-1, 1, Ny)
    M = g.ops.D[0]
    mean_vals = [2.5 + 0.2 * y, 0.4 * (1 - y * y), 0.02 * y, 0.1 * (1 - y * y),
                 1.0 + 0.05 * y]
    W = np.zeros((5, g.dNz, g.Xw, Ny), dtype=complex)
    for f in range(5):
        c = (rng.standard_normal((g.dNz, g.Xw, Ny))
             + 1j * rng.standard_normal((g.dNz, g.Xw, Ny))) * amp
        g.zero_dealiasing(c)
        phys = g.to_physical(c)              # real field, Hermitian by build
        w = g.to_wave(phys) / (g.dNx * g.dNz)
        g.zero_dealiasing(w)
        w[0, 0] = mean_vals[f]
        W[f] = np.linalg.solve(M, w.reshape(-1, Ny).T).T.reshape(w.shape)
    return W


def test_state_is_physical(grid):
    W = random_state(grid)
    s, _ = channel.aux_fields(grid, W)
    P = grid.to_physical(s)
    assert (P[4] > 0.8).all()
    refs = channel.references(grid, SCEN, P)
    assert (refs["T"] > 0).all()


def test_none_minus_rhome_is_implicit_operator(grid):
    W = random_state(grid)
    Nn, refs, _ = channel.apply_ns(grid, W, SCEN, "none")
    Nr, refs2, _ = channel.apply_ns(grid, W, SCEN, "rhome_xyz", refs=refs)
    lhs = Nn - Nr
    R = oracle.ns.imexop_refs(refs)
    sc = [SCEN[k] for k in ("Re", "Pr", "Ma", "alpha", "gamma")]
    scale = grid.dNx * grid.dNz
    worst = 0.0
    norm = np.abs(Nn).max()
    for iz in range(grid.dNz):
        for ix in range(grid.Xw):
            if not grid.keep[iz, ix]:
                assert not lhs[:, iz, ix].any()
                continue
            x = np.ascontiguousarray(W[:, iz, ix, :])
            km, kn = grid.kx[ix], grid.kz[iz]
            full = oracle.imexop_accumulate(1.0, km, kn, sc, R, grid.ops, x)
            mass = oracle.imexop_accumulate(0.0, km, kn, sc, R, grid.ops, x)
            Lx = (full - mass) * scale
            worst = max(worst, np.abs(lhs[:, iz, ix, :] - Lx).max())
    assert worst < 1e-11 * norm, (worst, norm)


def test_rhome_density_equation_is_zero(grid):
    W = random_state(grid)
    Nr, _, _ = channel.apply_ns(grid, W, SCEN, "rhome_xyz")
    assert not Nr[4].any()


def test_uniform_state_has_zero_rhs(grid):
    """A uniform state at rest is an equilibrium of the full equations."""
    W = np.zeros((5, grid.dNz, grid.Xw, grid.Ny), dtype=complex)
    M = grid.ops.D[0]
    ones = np.linalg.solve(M, np.ones(grid.Ny))
    W[0, 0, 0] = 2.5 * ones
    W[4, 0, 0] = 1.0 * ones
    Nn, _, _ = channel.apply_ns(grid, W, SCEN, "none")
    assert np.abs(Nn).max() < 1e-12 * grid.dNx * grid.dNz


def test_dt_candidates(grid):
    """Convective candidates scale as 1/|u| and diffusive ones are only
    formed where the local viscosity exceeds the reference
    (navier_stokes.hpp:766-892)."""
    W = random_state(grid)
    y = np.linspace(-1, 1, grid.Ny)
    lam = channel.lambdas(grid, y, 6, 2.0, 1.0, 1.0)
    _, refs, c = channel.apply_ns(grid, W, SCEN, "rhome_xyz", lam=lam)
    assert c.shape == (12,)
    assert (c > 0).all()
    # total convective bound cannot exceed any single-direction bound
    assert c[0] <= min(c[3], c[6], c[9]) + 1e-15
    _, _, c2 = channel.apply_ns(grid, 2 * W, SCEN, "rhome_xyz", lam=lam)
    assert c2[0] > 0


def _uniform_moving(Ny, nz, nx, rho0, vel, T0):
    """Physical fields of a uniform state with velocity vel and temperature
    T0 (p = rho T / gamma, e = p / (gamma - 1) + Ma^2 rho |u|^2 / 2) and
    vanishing derivatives."""
    gamma, Ma = SCEN["gamma"], SCEN["Ma"]
    P = np.empty((5, Ny, nz, nx))
    p = rho0 * T0 / gamma
    P[0] = p / (gamma - 1) + Ma * Ma / 2 * rho0 * sum(v * v for v in vel)
    for f, v in zip((1, 2, 3), vel):
        P[f] = rho0 * v
    P[4] = rho0
    A = {k: np.zeros((Ny, nz, nx)) for k in channel.AUX}
    return P, A


@pytest.mark.parametrize("lin", ["none", "rhome_xyz"])
def test_dt_candidates_closed_form_uniform_flow(lin):
    """navier_stokes.hpp:766-892 on a uniform moving state: every point has
    the same u, T and nu = T^beta / rho, so the 12 candidates reduce to
    closed forms in the extreme wall-normal scales.  Distinct U, V, W and
    distinct x, y, z scales catch a swapped direction, a missing sound
    speed or an evmaxmag mix-up."""
    Ny, nz, nx = 7, 4, 6
    rho0, vel, T0 = 1.3, (0.3, -0.2, 0.11), 1.7
    P, A = _uniform_moving(Ny, nz, nx, rho0, vel, T0)
    refs = channel.references(None, SCEN, P)
    l1x, l1z = 3.0, 5.0
    l1y = np.linspace(2.0, 11.0, Ny)[::-1].copy()
    l2y = np.linspace(4.0, 130.0, Ny)
    lam = (l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y)
    evr, evi = channel.SMR91_EVMAXMAG_REAL, channel.SMR91_EVMAXMAG_IMAG
    _, c = channel.rhs_physical(SCEN, P, A, refs, lin, lam, (evr, evi))
    a = math.sqrt(T0) / SCEN["Ma"] if lin == "none" else 0.0
    U, V, W = (abs(v) + a for v in vel)
    want = np.empty(12)
    tx, ty, tz = U * l1x, V * l1y.max(), W * l1z
    want[0] = evi / (tx + ty + tz)
    want[3], want[6], want[9] = evi / tx, evi / ty, evi / tz
    nu = T0 ** SCEN["beta"] / rho0
    md = (1 / SCEN["Re"]) * max(SCEN["gamma"] / SCEN["Pr"], max(1.0, SCEN["alpha"]))
    dx_, dy_, dz_ = nu * md * l1x * l1x, nu * md * l2y.max(), nu * md * l1z * l1z
    want[2] = evr / (nu * md * (l1x * l1x + l2y.max() + l1z * l1z))
    want[5], want[8], want[11] = evr / dx_, evr / dy_, evr / dz_
    conv = [0, 3, 6, 9]
    np.testing.assert_allclose(c[conv], want[conv], rtol=1e-14)
    if lin == "none":
        # no reference: fluctuating candidates equal the total ones and the
        # whole viscosity is diffusive
        np.testing.assert_allclose(c[[1, 4, 7, 10]], want[conv], rtol=1e-14)
        np.testing.assert_allclose(c[[2, 5, 8, 11]], want[[2, 5, 8, 11]],
                                   rtol=1e-14)
    else:
        # the reference profiles equal the state: the fluctuating velocity
        # and the excess viscosity vanish (to rounding), leaving the
        # candidates at or near their initial maximum
        assert (c[[1, 4, 7, 10]] > 1e12).all(), c
        assert (c[[2, 5, 8, 11]] > 1e12).all(), c


def test_smr91_step_keeps_rest_state(grid):
    """A uniform state at rest with T = 1 and isothermal T = 1 walls is a
    steady solution; one SMR91 step must leave it unchanged."""
    g = grid
    gamma = SCEN["gamma"]
    M = g.ops.D[0]
    ones = np.linalg.solve(M, np.ones(g.Ny))
    W = np.zeros((5, g.dNz, g.Xw, g.Ny), dtype=complex)
    W[0, 0, 0] = ones / (gamma * (gamma - 1))       # rho E at rho = T = 1
    W[4, 0, 0] = ones
    y = np.linspace(0, 2, g.Ny)
    lam = channel.lambdas(g, y, 6, 2.0, 1.0, 1.0)
    bc = ((1.0, 0, 0, 0), (1.0, 0, 0, 0))
    a, dt, _ = channel.smr91_step(g, W, SCEN, bc, lam, max_dt=1e-3)
    assert dt == 1e-3
    np.testing.assert_allclose(a, W, rtol=0, atol=1e-13)


def test_smr91_step_advances_perturbed_state(grid):
    g = grid
    W = random_state(g)
    y = np.linspace(0, 2, g.Ny)
    lam = channel.lambdas(g, y, 6, 2.0, 1.0, 1.0)
    bc = ((1.0, 0, 0, 0), (1.0, 0, 0, 0))
    a, dt, refs = channel.smr91_step(g, W, SCEN, bc, lam, max_dt=1e-3)
    assert np.isfinite(a).all()
    # every wavenumber meets the isothermal wall constraints afterwards:
    # rho E = E_factor rho and m = 0 at both walls (end coefficients)
    gamma = SCEN["gamma"]
    Ef = 1.0 / (gamma * (gamma - 1))
    scale = np.abs(a).max()
    for y in (0, g.Ny - 1):
        assert np.abs(a[0, :, :, y] - Ef * a[4, :, :, y]).max() < 1e-12 * scale
        assert np.abs(a[1:4, :, :, y]).max() < 1e-12 * scale
