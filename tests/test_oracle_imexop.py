"""Oracle pins for the compressible implicit operator M + phi L.

Follows tests/test_rholut_imexop.cpp:86-302 (operator consistency between
matrix-free accumulation and banded assembly, and (M+phi L)^{-1}(M+phi L) = I
through a banded LU) and adds an independent check of the term table: at a
uniform reference state at rest the operator must equal the textbook
linearized compressible Navier-Stokes operator (acoustics, Stokes flow with
bulk viscosity, heat conduction), derived here without the term table.
"""
import math

import numpy as np
import pytest

import oracle
from oracle.bspline import collocation, dense_of

# tests/test_rholut_imexop.cpp:96-118
BREAKPTS = [0.0, 0.5, 1.5, 2.25, 2.5, 3.0, 3.35]
PHI = complex(math.sqrt(2) / 11, math.log2(math.e) / 13)
SCEN = [3000.0, 0.7, 1.5, 3.0, 1.4]          # Re, Pr, Ma, alpha, gamma
# tests/test_rholut_imexop.cpp:320-330
WAVENUMBERS = [(0.0, 0.0), (0.0, 3 * math.pi), (7 * math.e, 0.0),
               (1.0, 3 * math.pi), (7 * math.e, 1.0), (7 * math.e, 3 * math.pi)]
S = 5


@pytest.fixture(scope="module")
def ops():
    return oracle.Ops(collocation(6, BREAKPTS, 2))


def refs_single(n, refndx):
    refs = np.zeros((oracle.NREF, n))
    if refndx >= 0:
        refs[refndx] = 1.0
    return refs


def dense_accumulate(ops, km, kn, refs, phi=PHI, scen=SCEN):
    n = ops.n
    N = S * n
    A = np.zeros((N, N), dtype=complex)
    for j in range(N):
        x = np.zeros((S, n), dtype=complex)
        x.reshape(-1)[j] = 1.0
        A[:, j] = oracle.imexop_accumulate(phi, km, kn, scen, refs, ops,
                                           x).reshape(-1)
    return A


def geometry(ops):
    g = {"KL": S * (ops.max_kl + 1) - 1, "KU": S * (ops.max_ku + 1) - 1}
    g["LD"] = g["KL"] + 1 + g["KU"]
    return g


def perm(n):
    """P as an index map: (P x)[i] = x[q(i)], q(i) = (i mod S) n + i / S."""
    i = np.arange(S * n)
    return (i % S) * n + i // S


def unpermute(papt_dense, n):
    """A^T from P A^T P^T."""
    q = perm(n)
    AT = np.zeros_like(papt_dense)
    AT[np.ix_(q, q)] = papt_dense
    return AT


def packed_dense(ops, km, kn, refs, packf=False, phi=PHI):
    g = geometry(ops)
    N = S * ops.n
    ld = g["LD"] + (g["KL"] if packf else 0)
    off = g["KL"] if packf else 0
    papt = np.full(N * ld, np.nan, dtype=complex)
    oracle.imexop_pack(phi, km, kn, SCEN, refs, ops, papt, ld, offset=off)
    band = papt.reshape(N, ld)[:, off:off + g["LD"]]
    return band, dense_of(band, N, g["KL"], g["KU"]), papt, ld


@pytest.mark.parametrize("km,kn", WAVENUMBERS)
@pytest.mark.parametrize("refndx", [-1] + list(range(26)))
def test_pack_matches_accumulate(ops, km, kn, refndx):
    refs = refs_single(ops.n, refndx)
    A = dense_accumulate(ops, km, kn, refs)
    band, PATPT, _, _ = packed_dense(ops, km, kn, refs)
    g = geometry(ops)
    N = S * ops.n
    for c in range(N):                        # every in-matrix band slot set
        r = np.arange(max(0, c - g["KU"]), min(N, c + g["KL"] + 1))
        assert not np.isnan(band[c, g["KU"] + r - c]).any()
    AT = unpermute(PATPT, ops.n)
    np.testing.assert_array_equal(AT.T, A)


@pytest.mark.parametrize("km,kn", WAVENUMBERS[::2] + [WAVENUMBERS[-1]])
def test_packf_equals_packc(ops, km, kn):
    refs = np.random.default_rng(5).uniform(-1, 1, (oracle.NREF, ops.n))
    bc, _, _, _ = packed_dense(ops, km, kn, refs, packf=False)
    bf, _, _, _ = packed_dense(ops, km, kn, refs, packf=True)
    np.testing.assert_array_equal(bc, bf)


@pytest.mark.parametrize("km,kn", WAVENUMBERS)
@pytest.mark.parametrize("refndx", [-1, 0, 4, 10, 14, 21, 24, 25])
def test_inverse_times_operator_is_identity(ops, km, kn, refndx):
    """tests/test_rholut_imexop.cpp:86-302 with the oracle's ZGBTRF/ZGBTRS
    in place of ZGBSVX: solve (P A^T P^T)^T y = P (A I) column by column."""
    n = ops.n
    N = S * n
    refs = refs_single(n, refndx)
    A = dense_accumulate(ops, km, kn, refs)
    g = geometry(ops)
    _, _, papt, ld = packed_dense(ops, km, kn, refs, packf=True)
    info, ipiv = oracle.zgbtrf(N, g["KL"], g["KU"], papt, ld)
    assert info == 0
    q = perm(n)
    B = np.ascontiguousarray(A[q, :].T)       # rows = right hand sides P A e_j
    oracle.zgbtrs("T", N, g["KL"], g["KU"], papt, ld, ipiv, B, nrhs=N, ldb=N)
    X = np.zeros((N, N), dtype=complex)
    X[q, :] = B.T                             # P^T y
    err = np.abs(X - np.eye(N)).max()
    cond = np.linalg.cond(A, 1)
    assert err < np.finfo(float).eps * cond, (err, cond)


def laplacian_blocks(ops, km, kn):
    M, D1, D2 = ops.D
    return M, D1, D2, D2 - (km * km + kn * kn) * M


@pytest.mark.parametrize("km,kn", [(0.0, 0.0), (7 * math.e, 3 * math.pi),
                                   (-2.5, 1.25)])
def test_operator_at_rest_is_textbook_linearization(ops, km, kn):
    """Uniform state at rest (rho0, e0, u = 0, kinematic viscosity nu0):
        d rho/dt = -div m
        d m/dt   = -(gamma-1)/Ma^2 grad e + nu0/Re (lap m + (alpha+1/3) grad div m)
        d e/dt   = -h0 div m + gamma nu0/(Re Pr) (lap e - (e0/rho0) lap rho)
    with h0 = (e0 + p0)/rho0, p0 = (gamma-1) e0, grad = (i km, d/dy, i kn)
    acting on B-spline coefficients through M and D1, lap = D2 - k^2 M."""
    Re, Pr, Ma, alpha, gamma = SCEN
    rho0, e0, nu0 = 1.3, 2.1, 0.7
    p0 = (gamma - 1) * e0
    n = ops.n
    refs = np.zeros((oracle.NREF, n))
    names = oracle.REF_NAMES
    refs[names.index("nu")] = nu0
    refs[names.index("e_divm")] = (e0 + p0) / rho0          # rholt.hpp:675-681
    mu0 = nu0 * rho0                                         # rholt.hpp:1481-1489
    refs[names.index("e_deltarho")] = mu0 / rho0**2 * ((gamma - 1) * e0 - 2 * p0)
    A = dense_accumulate(ops, km, kn, refs)

    M, D1, D2, lap = laplacian_blocks(ops, km, kn)
    Z = np.zeros_like(M)
    grad = [1j * km * M, D1, 1j * kn * M]
    E, U, V, W, R = range(5)
    L = [[Z.astype(complex) for _ in range(5)] for _ in range(5)]
    h0 = (e0 + p0) / rho0
    for a in range(3):
        L[R][U + a] = L[R][U + a] - grad[a]
        L[U + a][E] = L[U + a][E] - (gamma - 1) / Ma**2 * grad[a]
        L[E][U + a] = L[E][U + a] - h0 * grad[a]
        L[U + a][U + a] = L[U + a][U + a] + nu0 / Re * lap
    ap13 = alpha + 1.0 / 3.0
    # grad div m: component a gets d_a d_b m_b; products of first derivatives
    # along y use D2 (collocation of the second derivative), across use grads
    d = [1j * km, None, 1j * kn]
    for a in range(3):
        for b in range(3):
            if a == 1 and b == 1:
                term = D2
            elif a == 1:
                term = d[b] * D1
            elif b == 1:
                term = d[a] * D1
            else:
                term = d[a] * d[b] * M
            L[U + a][U + b] = L[U + a][U + b] + nu0 / Re * ap13 * term
    L[E][E] = L[E][E] + gamma * nu0 / (Re * Pr) * lap
    L[E][R] = L[E][R] - gamma * nu0 / (Re * Pr) * (e0 / rho0) * lap
    Lfull = np.block(L)
    Mfull = np.kron(np.eye(5), M)
    expect = Mfull + PHI * Lfull
    scale = np.abs(expect).max()
    np.testing.assert_allclose(A, expect, rtol=0, atol=1e-13 * scale)


def test_term_table_shape():
    """96 terms in 25 blocks; the density equation only sees momentum and
    never a reference profile (rholut_imexop.c:461-487)."""
    eq, inp, op, ref, k = oracle.imexop_terms(SCEN)
    assert len(eq) == 96
    rho_rows = eq == 4
    assert set(inp[rho_rows]) == {1, 2, 3}
    assert (ref[rho_rows] == -1).all()


def test_isothermal_enforcer(ops):
    """IsothermalPATPTEnforcer (operator_hybrid_isothermal.cpp:396-526):
    after op() + rhs() the solved state meets the wall constraints, and the
    interior equations are untouched."""
    rng = np.random.default_rng(11)
    n = ops.n
    N = S * n
    g = geometry(ops)
    refs = rng.uniform(0.5, 1.5, (oracle.NREF, n))
    km, kn = 7 * math.e, 3 * math.pi
    _, _, papt, ld = packed_dense(ops, km, kn, refs, packf=True)
    band0 = papt.reshape(N, ld)[:, g["KL"]:].copy()
    lower = (1.0, 0.0, 0.0, 0.0)
    upper = (1.2, 0.3, 0.0, -0.1)
    oracle.isothermal_op(S, n, g["KL"], g["KU"], papt, ld, SCEN, lower, upper,
                         offset=g["KL"])
    band1 = papt.reshape(N, ld)[:, g["KL"]:]
    # untouched columns are identical
    q = perm(n)
    walls = {0, n - 1}
    changed = {c for c in range(N)
               if not np.array_equal(band0[c], band1[c], equal_nan=True)}
    expect_changed = {c for c in range(N)
                      if q[c] // n in (0, 1, 2, 3) and q[c] % n in walls}
    assert changed == expect_changed
    # solve with a random rhs and check the constraints on the solution
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    pb = b[q].copy()
    oracle.isothermal_rhs(S, n, pb)
    info, ipiv = oracle.zgbtrf(N, g["KL"], g["KU"], papt, ld)
    assert info == 0
    oracle.zgbtrs("T", N, g["KL"], g["KU"], papt, ld, ipiv, pb)
    x = np.zeros(N, dtype=complex)
    x[q] = pb
    xs = x.reshape(S, n)
    # B-spline coefficients at the walls equal the wall values (collocation
    # at the end knots interpolates the end coefficients)
    Re, Pr, Ma, alpha, gamma = SCEN
    for y, (T, u, v, w) in ((0, lower), (n - 1, upper)):
        Ef = T / (gamma * (gamma - 1)) + Ma * Ma / 2 * (u * u + v * v + w * w)
        rho = xs[4, y]
        assert abs(xs[0, y] - Ef * rho) < 1e-12 * max(1, abs(rho))
        for k, vel in zip((1, 2, 3), (u, v, w)):
            assert abs(xs[k, y] - vel * rho) < 1e-12 * max(1, abs(rho))
    # interior equations still hold: rows of A for interior points
    A = dense_accumulate(ops, km, kn, refs)
    interior = [i * n + y for i in range(S) for y in range(n)
                if i == 4 or y not in walls]
    np.testing.assert_allclose(A[interior] @ x, b[interior], rtol=0,
                               atol=1e-10 * np.abs(b).max())
