"""Pins the oracle's bsmbsm permutation, pack and banded LU restatement.

Checks, each chosen so a plausible mistake fails one of them:
* q / q^{-1} against the Octave tables of tests/test_bsmbsm.cpp:48-152 and
  the identity q^{-1}(q(i)) = i (:186-198);
* zaPxpby against numpy fancy indexing (the permute+axpby equivalence of
  tests/test_bsmbsm.cpp:203-476) for both transposes, strides and scalars;
* pack + ZGBTRF + ZGBTRS against the 30-digit Mathematica solution of
  tests/test_bsmbsm.cpp:683-966 (real op and i*op) at 1.8e4 eps;
* ZGBTRF/ZGBTRS against scipy's LAPACK (an independent implementation):
  identical pivots, factors and solutions to rounding;
* the banded solve against a dense numpy solve of the de-permuted matrix.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle
from conftest import golden

EPS = np.finfo(float).eps


def test_q_qinv_octave_tables():
    g = golden("bsmbsm_golden.json")
    S, n = 5, 9
    assert [oracle.q(S, n, i) for i in range(S * n)] == g["q_S5n9"]
    assert [oracle.qinv(S, n, i) for i in range(S * n)] == g["qinv_S5n9"]


def test_q_identity_relation():
    for S in range(1, 5):
        for n in range(1, 10):
            for i in range(S * n):
                assert oracle.qinv(S, n, oracle.q(S, n, i)) == i


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("S,n", [(1, 1), (2, 7), (5, 10), (5, 96)])
@pytest.mark.parametrize("alpha,beta", [(1, 0), (1, 1), (2.5, -0.5),
                                        (1 + 2j, 3 - 1j)])
def test_zaPxpby_matches_fancy_indexing(trans, S, n, alpha, beta):
    rng = np.random.default_rng(S * 100 + n)
    N = S * n
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y0 = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = y0.copy()
    oracle.zaPxpby(trans, S, n, alpha, x, beta, y)
    i = np.arange(N)
    perm = (i % S) * n + i // S if trans == "N" else (i % n) * S + i // n
    np.testing.assert_allclose(y, alpha * x[perm] + beta * y0, rtol=1e-15,
                               atol=1e-15)


def test_zaPxpby_negative_strides():
    S, n = 3, 4
    N = S * n
    x = np.arange(N, dtype=np.complex128) * (1 + 1j)
    y = np.zeros(N, dtype=np.complex128)
    oracle.zaPxpby("N", S, n, 1, x[::-1].copy(), 0, y, incx=-1)
    i = np.arange(N)
    np.testing.assert_array_equal(y, x[(i % S) * n + i // S])


def _golden_operator():
    g = golden("bsmbsm_golden.json")["solve"]
    f = lambda v: np.array([float(Fraction(s)) for s in v])
    return g, f(g["M"]), f(g["D1"]), f(g["D2"]), f(g["BR"]), f(g["XR"])


def _pack_golden(scale):
    """Build PAP^T of op (or i*op) in LU-ready storage with oracle.zpack."""
    g, M, D1, D2, BR, XR = _golden_operator()
    S, n, kl, ku = g["S"], g["n"], g["kl"], g["ku"]
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    LD = KL + 1 + KU
    ldab = LD + KL
    ab = np.full(n * S * ldab, np.nan, dtype=np.complex128)
    view = ab[KL:]          # packf: offset KL, leading dimension LD + KL
    blocks = {(0, 0): (1, M), (1, 1): (2, M), (2, 2): (4, M),
              (0, 1): (1 / 5, D1), (2, 1): (1 / 10, D1),
              (1, 0): (1 / 7, D2), (1, 2): (1 / 14, D2),
              (0, 2): (0, None), (2, 0): (0, None)}
    for (ih, jh), (a, B) in blocks.items():
        b = None if B is None else B.astype(np.complex128)
        oracle.zpack(S, n, ih, jh, kl, ku, a * scale, b, kl + 1 + ku, KL, KU,
                     view, ldab)
    return g, ab, ldab, KL, KU, BR, XR


@pytest.mark.parametrize("scale", [1, 1j])
def test_golden_solve_30_digits(scale):
    g, ab, ldab, KL, KU, BR, XR = _pack_golden(scale)
    S, n = g["S"], g["n"]
    N = S * n
    # no NaN left inside the band rows of the LU-ready storage
    band = ab.reshape(N, ldab)[:, KL:]
    for c in range(N):
        for r in range(max(0, c - KU), min(N, c + KL + 1)):
            assert not np.isnan(band[c, KU + r - c]), (r, c)
    rhs = np.zeros(N, dtype=np.complex128)
    oracle.zaPxpby("N", S, n, 1, (BR * scale).astype(np.complex128), 0, rhs)
    info, ipiv = oracle.zgbtrf(N, KL, KU, ab, ldab)
    assert info == 0
    oracle.zgbtrs("N", N, KL, KU, ab, ldab, ipiv, rhs)
    x = np.zeros(N, dtype=np.complex128)
    oracle.zaPxpby("T", S, n, 1, rhs, 0, x)
    tol = EPS * g["tolerance_eps_multiple"]
    np.testing.assert_allclose(x.real, XR, rtol=tol, atol=tol * np.abs(XR).max())
    np.testing.assert_allclose(x.imag, 0, atol=tol * np.abs(XR).max())


def _random_band(rng, N, kl, ku, dtype=np.complex128):
    ldab = 2 * kl + ku + 1
    ab = np.zeros((N, ldab), dtype=dtype)
    A = np.zeros((N, N), dtype=dtype)
    for j in range(N):
        for i in range(max(0, j - ku), min(N, j + kl + 1)):
            v = rng.standard_normal() + 1j * rng.standard_normal()
            ab[j, kl + ku + i - j] = v
            A[i, j] = v
    return ab, A


@pytest.mark.parametrize("N,kl,ku", [(1, 0, 0), (7, 2, 3), (40, 9, 4),
                                     (120, 14, 14)])
def test_gbtrf_gbtrs_match_scipy_lapack(N, kl, ku):
    rng = np.random.default_rng(N + kl)
    ab, A = _random_band(rng, N, kl, ku)
    ldab = ab.shape[1]
    lub, piv, info_s = la.zgbtrf(ab.T.copy(), kl, ku)
    mine = ab.reshape(-1).copy()
    info, ipiv = oracle.zgbtrf(N, kl, ku, mine, ldab)
    assert info == info_s == 0
    np.testing.assert_array_equal(ipiv, piv + 1)
    got = mine.reshape(N, ldab).T
    # compare U (rows 0..kl+ku) and multipliers where LAPACK defines them
    scale = np.abs(lub).max()
    np.testing.assert_allclose(got[kl:], lub[kl:], rtol=0, atol=1e-13 * scale)
    for trans in ("N", "T", "C"):
        b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        xs, info_t = la.zgbtrs(lub, kl, ku, b.copy(), piv,
                               trans={"N": 0, "T": 1, "C": 2}[trans])
        x = oracle.zgbtrs(trans, N, kl, ku, mine, ldab, ipiv, b.copy())
        np.testing.assert_allclose(x, xs, rtol=1e-11, atol=1e-12)
        opA = {"N": A, "T": A.T, "C": A.conj().T}[trans]
        np.testing.assert_allclose(opA @ x, b, rtol=1e-10, atol=1e-10)


def test_gbtrf_zero_pivot_reports_info():
    N, kl, ku = 6, 1, 1
    ab = np.zeros((N, 2 * kl + ku + 1), dtype=np.complex128)
    for j in range(N):
        ab[j, kl + ku] = 1.0         # identity ...
    ab[3, kl + ku] = 0.0             # ... with a hole on the diagonal
    flat = ab.reshape(-1).copy()
    info, _ = oracle.zgbtrf(N, kl, ku, flat, ab.shape[1])
    _, _, info_s = la.zgbtrf(ab.T.copy(), kl, ku)
    assert info == info_s == 4


def test_cgbtrf_cgbtrs_single_precision():
    rng = np.random.default_rng(3)
    N, kl, ku = 50, 6, 5
    ab, A = _random_band(rng, N, kl, ku)
    abs_ = ab.astype(np.complex64).reshape(-1).copy()
    info, ipiv = oracle.zgbtrf(N, kl, ku, abs_, ab.shape[1])
    lub, piv, _ = la.cgbtrf(ab.T.astype(np.complex64).copy(), kl, ku)
    assert info == 0
    np.testing.assert_array_equal(ipiv, piv + 1)
    b = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    x = oracle.zgbtrs("N", N, kl, ku, abs_, ab.shape[1], ipiv, b.copy())
    np.testing.assert_allclose(A @ x.astype(np.complex128), b, rtol=1e-4,
                               atol=1e-4)


def test_pack_matches_dense_permutation():
    """PAP^T from 4 packs equals A[q][:, q] of the dense block matrix."""
    rng = np.random.default_rng(7)
    S, n, kl, ku = 2, 6, 2, 1
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    LD = KL + 1 + KU
    N = S * n
    A = np.zeros((N, N), dtype=np.complex128)
    papt = np.full(N * LD, np.nan, dtype=np.complex128)
    for ih in range(S):
        for jh in range(S):
            B = np.zeros((n, n), dtype=np.complex128)
            for j in range(n):
                for i in range(max(0, j - ku), min(n, j + kl + 1)):
                    B[i, j] = rng.standard_normal() + 1j * rng.standard_normal()
            A[ih * n:(ih + 1) * n, jh * n:(jh + 1) * n] = B
            bb = np.zeros((n, kl + 1 + ku), dtype=np.complex128)
            for j in range(n):
                for i in range(max(0, j - ku), min(n, j + kl + 1)):
                    bb[j, ku + i - j] = B[i, j]
            oracle.zpack(S, n, ih, jh, kl, ku, 1, bb.reshape(-1), kl + 1 + ku,
                         KL, KU, papt, LD)
    i = np.arange(N)
    qv = (i % S) * n + i // S
    PAPT = A[qv][:, qv]
    band = papt.reshape(N, LD)
    for c in range(N):
        for r in range(max(0, c - KU), min(N, c + KL + 1)):
            assert band[c, KU + r - c] == PAPT[r, c]


def test_zcgbsvx_lotkin_known_answer():
    """Lotkin (1955) matrices with x = 1 (tests/test_lapackext.c:42-71)."""
    for n in range(1, 9):
        kl = ku = n - 1
        ldab, ldafb = kl + 1 + ku, 2 * kl + 1 + ku
        ab = np.zeros((n, ldab), dtype=np.complex128)
        b = np.zeros(n, dtype=np.complex128)
        for j in range(n):
            ab[j, ku + 0 - j] = 1
        b[0] = n
        for i in range(1, n):
            s = Fraction(0)
            for j in range(n - 1, -1, -1):
                a = Fraction(1, i + j + 1)
                s += a
                ab[j, ku + i - j] = float(a)
            b[i] = float(s)
        afb = np.zeros(n * ldafb, dtype=np.complex128)
        x = np.zeros(n, dtype=np.complex128)
        ipiv = np.zeros(n, dtype=np.int32)
        out = oracle.zcgbsvx("N", 0, 5, "N", n, kl, ku, ab.reshape(-1).copy(),
                             afb, ipiv, b.copy(), x, siter=25, diter=10,
                             tolsc=1.0)
        assert out["info"] == 0
        cond_scale = 10.0 ** (1.5 * n)     # Lotkin conditioning grows fast
        np.testing.assert_allclose(x.real, 1, rtol=0, atol=1e-15 * cond_scale)
        assert out["fact"] in ("S", "D")
