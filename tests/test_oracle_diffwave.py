"""Pin for the oracle's diffwave (oracle/oracle.c sz_or_diffwave_accumulate):
the integer-field closed form of tests/test_diffwave.cpp:55-406.

The fields are x(l, m, n) = (l+3)(m+4)(n+6) - i (l+8)(m+12)(n+14) and
y(l, m, n) = (l+18)(m+20)(n+24) - i (l+30)(m+32)(n+38) on the (Y, X, Z)
column-major wave layout, alpha = 2, beta = 3 (accumulate) and alpha = 2+1i
or 1 (apply).  Mode (m, n) is kept when both |w| <= (N-1)/2 (in-order
wavenumber w of the dealiased index); a kept mode gets
alpha (i 2pi/Lx w_m)^dx (i 2pi/Lz w_n)^dz x + beta y, the others beta y
(apply: zero).  The expectation is formed here with Python complex powers,
independently of the oracle's square-and-multiply, over dx, dz in 0..4 and
the sixteen extents of the reference test.  The reference's absolute
tolerance 7 10^(-10 + (dx+dz)/2.5) is kept, and a relative 1e-13 bound is
added (the two evaluations differ only in rounding).
"""
import math

import numpy as np
import pytest

import oracle

CASES = [  # Lx, Lz, Ny, Nx, dNx, Nz, dNz (test_diffwave.cpp:184-203)
    (5, 7, 1, 1, 1, 6, 9), (5, 7, 1, 1, 1, 6, 8), (5, 7, 1, 1, 1, 5, 8),
    (5, 7, 1, 1, 1, 5, 9), (5, 7, 1, 6, 9, 1, 1), (5, 7, 1, 6, 8, 1, 1),
    (5, 7, 1, 5, 8, 1, 1), (5, 7, 1, 5, 9, 1, 1), (5, 7, 1, 6, 9, 6, 9),
    (5, 7, 1, 6, 8, 6, 8), (5, 7, 1, 5, 8, 5, 8), (5, 7, 1, 5, 9, 5, 9),
    (5, 7, 3, 4, 4, 4, 4), (5, 7, 3, 8, 8, 8, 8), (5, 7, 3, 7, 7, 7, 7),
    (5, 7, 3, 8, 12, 16, 24)]


def _fields(Ny, dNx, dNz):
    X = dNx // 2 + 1
    n, m, l = np.meshgrid(np.arange(dNz), np.arange(X), np.arange(Ny),
                          indexing="ij")
    x = (l + 3.0) * (m + 4) * (n + 6) - 1j * ((l + 8.0) * (m + 12) * (n + 14))
    y = (l + 18.0) * (m + 20) * (n + 24) - 1j * ((l + 30.0) * (m + 32) * (n + 38))
    return x, y


def _inorder_abs(N, i):
    return i if i < N // 2 + 1 else N - i


def _inorder_diff(N, dN, i):
    if i < (N + 1) // 2:
        return i
    if i >= dN - (N - 1) // 2:
        return -dN + i
    return 0


def _expected(dx, dz, alpha, x, beta, y, Lx, Lz, Nx, dNx, Nz, dNz):
    out = beta * y if beta is not None else np.zeros_like(x)
    for n in range(x.shape[0]):
        kz = 1j * (2 * math.pi / Lz) * _inorder_diff(Nz, dNz, n)
        zs = 1.0 if dz == 0 else kz ** dz
        for m in range(x.shape[1]):
            kx = 1j * (2 * math.pi / Lx) * _inorder_diff(Nx, dNx, m)
            xs = 1.0 if dx == 0 else kx ** dx
            keep = (_inorder_abs(dNz, n) <= (Nz - 1) // 2
                    and _inorder_abs(dNx, m) <= (Nx - 1) // 2)
            if keep:
                out[n, m] = out[n, m] + (xs * zs) * alpha * x[n, m]
    return out


def _check(got, want, dx, dz):
    small = 7 * 10 ** (-10 + (dx + dz) / 2.5)
    err = np.abs(got - want)
    assert err.max() <= small, (err.max(), small)
    scale = np.maximum(np.abs(want), 1.0)
    assert (err / scale).max() <= 1e-13, (err / scale).max()


@pytest.mark.parametrize("dx", range(5))
@pytest.mark.parametrize("dz", range(5))
def test_diffwave_accumulate_closed_form(dx, dz):
    for Lx, Lz, Ny, Nx, dNx, Nz, dNz in CASES:
        x, y = _fields(Ny, dNx, dNz)
        got = oracle.diffwave_accumulate(dx, dz, 2.0, x.copy(), 3.0, y.copy(),
                                         Lx, Lz, Ny, Nx, dNx, Nz, dNz)
        want = _expected(dx, dz, 2.0, x, 3.0, y, Lx, Lz, Nx, dNx, Nz, dNz)
        _check(got, want, dx, dz)


@pytest.mark.parametrize("dx", range(5))
@pytest.mark.parametrize("dz", range(5))
def test_diffwave_apply_closed_form(dx, dz):
    for Lx, Lz, Ny, Nx, dNx, Nz, dNz in CASES:
        x, _ = _fields(Ny, dNx, dNz)
        for alpha in (2 + 1j, 1.0):
            got = oracle.diffwave_accumulate(dx, dz, alpha, x.copy(), 0.0,
                                             np.zeros_like(x), Lx, Lz, Ny, Nx,
                                             dNx, Nz, dNz)
            want = _expected(dx, dz, alpha, x, None, None, Lx, Lz, Nx, dNx,
                             Nz, dNz)
            _check(got, want, dx, dz)
            if dx == 0 and dz == 0 and alpha == 1.0:
                # no derivative and unit scaling: exact on kept modes
                np.testing.assert_array_equal(got, want)
