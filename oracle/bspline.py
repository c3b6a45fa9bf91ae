"""Greville collocation B-spline operators for the oracle (numpy + scipy).

TEST INFRASTRUCTURE ONLY.  Restates suzerain's collocation operator setup:

* breakpoints: suzerain/support/support.cpp:289-308 (linspace of ndof-k+2
  points on [0, 1] mapped by htstretch2 (htdelta >= 0) or htstretch1
  (htdelta < 0), suzerain/htstretch.c:40-53,112-124, then scaled to
  [left, right]); linspace as suzerain/math.hpp:194-216;
* knots: k-fold end knots, simple interior breakpoints (GSL's
  gsl_bspline_knots), n = nbreak + k - 2 basis functions;
* Greville abscissae: mean of knots i+1..i+k-1;
* D_d(i, j) = B_j^{(d)}(xi_i) (suzerain/bsplineop.c:581-652), evaluated here
  with scipy.interpolate.BSpline -- an implementation independent of the
  product's de Boor code;
* per-derivative bandwidths of the TRANSPOSED storage trimmed from k-1 while
  the outermost diagonal vanishes in both k x k corner blocks
  (suzerain/bsplineop.c:401-517).
"""
import math
from fractions import Fraction

import numpy as np
from scipy.interpolate import BSpline


def htstretch1(delta, L, x):
    if delta == 0:
        return x / L
    return 1 + math.tanh(delta * (x / L - 1)) / math.tanh(delta)


def htstretch2(delta, L, x):
    if delta == 0:
        return x / L
    return 0.5 * (1 + math.tanh(delta * (x / L - 0.5)) / math.tanh(delta / 2.0))


def breakpoints(ndof, k, left, right, htdelta):
    nb = ndof - k + 2
    xh = 1.0 / (nb - 1)
    u = np.array([0.0] + [xh * i + 0.0 for i in range(1, nb - 1)] + [1.0])
    if htdelta >= 0:
        s = np.array([htstretch2(+htdelta, 1.0, v) for v in u])
    else:
        s = np.array([htstretch1(-htdelta, 1.0, v) for v in u])
    return (right - left) * s + left


def knots(k, bp):
    bp = [float(b) for b in bp]
    return np.array([bp[0]] * k + bp[1:-1] + [bp[-1]] * k)


def greville(k, t):
    n = len(t) - k
    return np.array([np.mean(t[i + 1:i + k]) for i in range(n)])


def greville_exact(k, bp):
    """Greville abscissae as Fractions for rational breakpoints."""
    bp = [Fraction(b) for b in bp]
    t = [bp[0]] * k + bp[1:-1] + [bp[-1]] * k
    n = len(t) - k
    return [sum(t[i + 1:i + k]) / (k - 1) for i in range(n)]


def collocation(k, bp, nderiv):
    """Dense D[d] (n x n) for d = 0..nderiv plus trimmed transposed bands.

    Returns dict(D=[...], kl=[...], ku=[...], points=xi, n=n) where kl/ku are
    the bandwidths of the TRANSPOSED storage as in the reference.
    """
    t = knots(k, bp)
    n = len(t) - k
    xi = greville(k, t)
    xi[0], xi[-1] = t[0], t[-1]
    D = []
    for d in range(nderiv + 1):
        M = np.zeros((n, n))
        for j in range(n):
            c = np.zeros(n)
            c[j] = 1.0
            spl = BSpline(t, c, k - 1, extrapolate=False)
            if d > 0:
                spl = spl.derivative(d) if d <= k - 1 else None
            if spl is None:
                continue
            v = spl(xi)
            # right end point: evaluate the last interval's polynomial there
            v_end = BSpline(t, c, k - 1, extrapolate=True)
            if d > 0:
                v_end = v_end.derivative(d)
            v[-1] = v_end(xi[-1] - 0.0)
            M[:, j] = np.nan_to_num(v)
        D.append(M)
    kl, ku = [], []
    for M in D:
        kT = k - 1
        uT = k - 1
        # transposed superdiagonal s <-> D(i, i-s); subdiagonal s <-> D(i, i+s)
        def corner_sum(off):
            s = 0.0
            for lo in (0, n - k):
                for i in range(lo, lo + k):
                    j = i + off
                    if 0 <= i < n and lo <= j < lo + k and 0 <= j < n:
                        s += abs(M[i, j])
            return s
        for s_ in range(k - 1, 0, -1):
            if corner_sum(-s_) == 0.0:
                uT -= 1
            else:
                break
        for s_ in range(k - 1, 0, -1):
            if corner_sum(+s_) == 0.0:
                kT -= 1
            else:
                break
        kl.append(kT)
        ku.append(uT)
    return {"D": D, "kl": kl, "ku": ku, "points": xi, "n": n, "knots": t}


def integration_coefficients(k, bp):
    """Integral over [bp[0], bp[-1]] of each basis function
    (suzerain_bspline_integration_coefficients, suzerain/bspline.c:223-282,
    nderiv = 0): a (k+1)//2-point Gauss-Legendre rule on every breakpoint
    interval, the basis evaluated with scipy.interpolate.BSpline."""
    t = knots(k, bp)
    n = len(t) - k
    xg, wg = np.polynomial.legendre.leggauss((k + 1) // 2)
    out = np.zeros(n)
    for a, b in zip(bp[:-1], bp[1:]):
        x = (b - a) / 2 * xg + (a + b) / 2
        wts = (b - a) / 2 * wg
        for j in range(n):
            c = np.zeros(n)
            c[j] = 1.0
            v = np.nan_to_num(BSpline(t, c, k - 1, extrapolate=False)(x))
            out[j] += np.dot(wts, v)
    return out


def gram(k, bp):
    """Dense Gram matrix G(i, j) = integral of B_i B_j over [bp[0], bp[-1]]
    (the Galerkin L2 operator of bsplineop.c, nderiv = 0): a k-point
    Gauss-Legendre rule per breakpoint interval (exact for the degree 2k-2
    products), the basis evaluated with scipy.interpolate.BSpline."""
    t = knots(k, bp)
    n = len(t) - k
    xg, wg = np.polynomial.legendre.leggauss(k)
    G = np.zeros((n, n))
    for a, b in zip(bp[:-1], bp[1:]):
        x = (b - a) / 2 * xg + (a + b) / 2
        wts = (b - a) / 2 * wg
        B = np.zeros((n, len(x)))
        for j in range(n):
            c = np.zeros(n)
            c[j] = 1.0
            B[j] = np.nan_to_num(BSpline(t, c, k - 1, extrapolate=False)(x))
        G += (B * wts) @ B.T
    return G


def band_of(M, kl, ku):
    """Column-major general band storage (ld = kl+1+ku) of dense M, shaped
    (n, ld): ab[j, ku + i - j] = M[i, j] inside the band, zero elsewhere."""
    n = M.shape[0]
    ld = kl + 1 + ku
    j = np.arange(n)[:, None]
    r = np.arange(ld)[None, :]
    i = j + r - ku
    ok = (i >= 0) & (i < n)
    ab = np.zeros((n, ld), dtype=M.dtype)
    ab[ok] = M[i[ok], np.broadcast_to(j, i.shape)[ok]]
    return ab


def dense_of(ab, n, kl, ku):
    """Inverse of band_of for ab shaped (n, ld)."""
    M = np.zeros((n, n), dtype=ab.dtype)
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            M[i, j] = ab[j, ku + i - j]
    return M
